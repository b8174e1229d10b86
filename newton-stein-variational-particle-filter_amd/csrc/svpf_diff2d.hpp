// svpf_diff2d.hpp -- the two-factor (2-D) SVPF with checkpoints and its
// hand-written adjoint.
//
// Particle x = (f, s), log-vol h = mu + f + s (the reference's two-factor
// decomposition, src/svpf_two_factor.cu, include/svpf_two_factor.cuh), moved
// by one 2-D IMQ kernel with a 2x2 Newton preconditioner:
//   predict    m_i = (rho_f f_i, rho_s s_i), x_i <- clamp(m_i + (sigma_f z_f, sigma_s z_s), +-20)
//   bandwidth  ell_c^2 = lower median of the pairwise (x_ic - x_jc)^2, bw_c = clamp(ell_c, .01, 2)
//   score      g_j = clamp(g_prior_j + beta (1, 1) g_lik(h_j), +-10), the mixture prior
//              g_prior = -sum_i r_ij D (x_j - m_i) / sum_i r_ij, D = diag(sigma_f^-2, sigma_s^-2)
//   Newton     C_j = D + c_j 1 1^T (c_j = -d^2 log p(y|h)/dh^2), s_j = 0.7 C_j^-1 g_j, Q_j = +-2 C_j^-1
//   transport  phi_i = (1/N) sum_j K_ij s_j + K_ij^2 Q_j M (x_i - x_j),
//              K = 1/(1 + (x_i-x_j)^T M (x_i-x_j)), M = diag(bw_f^-2, bw_s^-2)
//   update     RMSProp and SVLD noise per coordinate.
// Positions are scaled per coordinate (x_c / bw_c in the transport,
// sqrt(log2 e / 2) (x_c - m_c) / sigma_c in the prior).  The structure follows
// svpf_diff.hpp: forward with checkpoints (or the whole tape when the
// interval covers T), backward by segments, two O(N^2) adjoint passes per
// Stein iteration:
//   pass C: for each row, w_i . df/dd summed over j (row part) and the sums
//     over j of K w_j, K^2 w_j d^T, K^2 w_j and K^3 d (w_j^T Qt_i d) (column part);
//   pass B: for each component i, sum over j of rt_ij (a_j - psi_ij kappa sq * dd_ij).
#pragma once

#include "svpf_device.hpp"
#include "svpf_diff.hpp"

namespace svpf {

constexpr float kD2Clamp = 20.0f;
// Rows of the optional per-iteration buffer svpf_d2desc.drift ([T][S][kD2DriftRows][n]): the Stein drift
// (phi_f, phi_s) and the mixture prior's shift (the nearest component's scaled squared distance), so that
// the backward's recompute replays positions without the transport and the prior without its
// nearest-component pass.
constexpr int kD2DriftRows = 3;

// Tape slot of one tick (floats): x [2n], u [2n], then per iteration the
// T2_ROWS rows of n floats, then 8 scalars (ell2_f, ell2_s, a_f, b_f, a_s, b_s).
// The prior's third moments are not taped: the sigma adjoint of the mixture
// prior follows from its homogeneity (see tick_back).
enum {
    T2_HF = 0, T2_HS, T2_VF, T2_VS, T2_PF, T2_PS, T2_UF, T2_US, T2_EF, T2_ES,
    T2_VFF, T2_VFS, T2_VSS, T2_L, T2_ROWS
};
__host__ __device__ inline size_t diff2d_slot_floats(int n, int S) { return (size_t)(4 + T2_ROWS * S) * n + 8; }

struct D2Theta {
    float mu, rho_f, sig_f, rho_s, sig_s, nu;
    float st_const, dst_const;
};

SVPF_DEV D2Theta d2_theta(const float* th) {
    D2Theta t;
    t.mu = th[0];
    t.rho_f = th[1];
    t.sig_f = th[2];
    t.rho_s = th[3];
    t.sig_s = th[4];
    t.nu = th[5];
    t.st_const = lgammaf((t.nu + 1.0f) * 0.5f) - lgammaf(t.nu * 0.5f) - 0.5f * logf(3.14159265358979f * t.nu);
    t.dst_const = 0.5f * svpf_digamma((t.nu + 1.0f) * 0.5f) - 0.5f * svpf_digamma(t.nu * 0.5f) - 0.5f / t.nu;
    return t;
}

struct D2Lds {
    __host__ __device__ static int rec_len(int nmax) { return nmax < 32 ? 32 : nmax; }
    __host__ __device__ static size_t bytes(int nmax) {
        size_t b = (size_t)rec_len(nmax) * 16 * 3 + (size_t)DiffLds::mbuf_len(nmax) * 8 + kRedWords * 4 + 16 * 4;
        return (b + 15) & ~(size_t)15;
    }
};

template <int BS, int R>
struct D2Engine {
    float4* rec;   // [nmax] {u_f, u_s, s_f, s_s}; pass B {hh_f, hh_s, L, b}
    float4* rec2;  // [nmax] {Qt_ff, Qt_fs, Qt_sf, Qt_ss}; pass B {at_f, at_s, a_f, a_s}
    float4* rec3;  // [nmax] {w_f, w_s, 0, 0}; pair-search scratch
    float* mbuf;   // [2 p2] scaled means (mm_f, mm_s) interleaved / sort buffer
    float* red;
    float* scal;
    int n, p2;
    float inv_n;
    float xf[R], xs[R], vf[R], vs[R];

    SVPF_DEV void bind(char* lds, int n_, int nmax) {
        n = __builtin_amdgcn_readfirstlane(n_);
        p2 = DiffLds::mbuf_len(n);
        const int rl = D2Lds::rec_len(nmax);
        rec = reinterpret_cast<float4*>(lds);
        rec2 = rec + rl;
        rec3 = rec2 + rl;
        mbuf = reinterpret_cast<float*>(rec3 + rl);
        red = mbuf + 2 * DiffLds::mbuf_len(nmax);
        scal = red + kRedWords;
        inv_n = 1.0f / (float)n;
    }
    SVPF_DEV static int row(int r) { return (int)threadIdx.x + r * BS; }
    SVPF_DEV bool valid(int r) const { return row(r) < n; }

    // predict normals (f, s) of particle i at tick t, and Langevin normals of iteration k
    SVPF_DEV void noise_z(const svpf_d2desc& D, int t, int i, float& zf, float& zs) const {
        if (D.z) {
            zf = D.z[((size_t)t * 2) * n + i];
            zs = D.z[((size_t)t * 2 + 1) * n + i];
            return;
        }
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        uint32_t o[4];
        rng.draw((uint32_t)i, (uint32_t)t, SVPF_PURPOSE_PREDICT, o);
        zf = normal_from_bits(o[0], o[1]);
        zs = normal_from_bits(o[2], o[3]);
    }
    SVPF_DEV void noise_xi(const svpf_d2desc& D, int S, int t, int k, int i, float& a, float& b) const {
        if (D.z) {
            a = D.xi[(((size_t)t * S + k) * 2) * n + i];
            b = D.xi[(((size_t)t * S + k) * 2 + 1) * n + i];
            return;
        }
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        uint32_t o[4];
        rng.draw((uint32_t)i, (uint32_t)t, SVPF_PURPOSE_SVLD0 + (uint32_t)k, o);
        a = normal_from_bits(o[0], o[1]);
        b = normal_from_bits(o[2], o[3]);
    }

    SVPF_DEV float logmeanexp(const float (&a)[R]) {
        float mx = -3.0e38f;
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) mx = fmaxf(mx, a[r]);
        mx = block_max<BS>(mx, red);
        float se[1] = {0.f};
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) se[0] += expf(a[r] - mx);
        block_sum<BS, 1>(se, red);
        return mx + logf(fmaxf(se[0] * inv_n, 1e-10f));
    }

    // lower median of the pairwise squared distances of one coordinate (sort + selection)
    SVPF_DEV float median_coord(const float (&h)[R], float est) {
        float* tmp = reinterpret_cast<float*>(rec);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) tmp[row(r)] = h[r];
        __syncthreads();
        block_sort<BS>(tmp, mbuf, n);
        int* plo = reinterpret_cast<int*>(rec);
        constexpr int KR = R < 2 ? 2 : R;
        float med = median_pairwise_r2<BS, KR>(mbuf, n, est, plo, plo + n, reinterpret_cast<float*>(plo + 2 * n),
                                               2 * n, reinterpret_cast<uint32_t*>(red));
        __syncthreads();
        return med;
    }

    // first pair (a, b), a < b, of one coordinate whose squared distance is ell2.
    // sorted: mbuf holds the coordinate's values ascending (its median's sort); otherwise they are sorted here
    SVPF_DEV void median_pair(const float (&h)[R], float ell2, bool sorted, int& a, int& b) {
        if (!sorted) {
            float* tmp = reinterpret_cast<float*>(rec);
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) tmp[row(r)] = h[r];
            __syncthreads();
            block_sort<BS>(tmp, mbuf, n);
        }
        float* hl = reinterpret_cast<float*>(rec3);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) hl[row(r)] = h[r];
        first_pair_at_r2<BS, R>(h, hl, mbuf, n, ell2, reinterpret_cast<int*>(hl + n), reinterpret_cast<int*>(scal + 8),
                                a, b);
    }

    // ---- forward tick ----
    template <bool RECORD, bool SELECT>
    SVPF_DEV float tick(const svpf_d2config& C, const D2Theta& th, const svpf_d2desc& D, int t, float* slot,
                        float (&med)[2], float (&outv)[8]) {
        const int S = C.S;
        const float y_t = D.y[t];
        const float ysq = y_t * y_t;
        const float nu = th.nu, np1 = th.nu + 1.0f;
        float mf[R], ms[R], hf[R], hs[R], lw[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            const bool ok = i < n;
            float zf = 0.f, zs = 0.f;
            if (ok) noise_z(D, t, i, zf, zs);
            mf[r] = th.rho_f * xf[r];
            ms[r] = th.rho_s * xs[r];
            const float uf = mf[r] + th.sig_f * zf, us = ms[r] + th.sig_s * zs;
            hf[r] = fminf(fmaxf(uf, -kD2Clamp), kD2Clamp);
            hs[r] = fminf(fmaxf(us, -kD2Clamp), kD2Clamp);
            if (RECORD && ok) {
                slot[i] = xf[r];
                slot[n + i] = xs[r];
                slot[2 * (size_t)n + i] = uf;
                slot[3 * (size_t)n + i] = us;
            }
            const float hh = th.mu + hf[r] + hs[r];
            const float A = ysq / (safe_expf(hh) + 1e-8f) / nu;
            lw[r] = ok ? th.st_const - 0.5f * hh - np1 * 0.5f * log1pf(A) : -3.0e38f;
        }
        const float pred_ll = logmeanexp(lw);
        float e2[2];
        int pr[4] = {-1, -1, -1, -1};
#pragma unroll
        for (int c = 0; c < 2; ++c) {
            // the pair search reads the coordinate's sorted values, so it follows its own median
            if (SELECT) {
                e2[c] = median_coord(c == 0 ? hf : hs, med[c]);
                med[c] = e2[c];
                if (threadIdx.x == 0) D.ell2[2 * t + c] = e2[c];
            } else {
                e2[c] = D.ell2[2 * t + c];
            }
            if (RECORD) {
                const float b_ = sqrtf(e2[c]);
                if (b_ > 0.01f && b_ < 2.0f) median_pair(c == 0 ? hf : hs, e2[c], SELECT, pr[2 * c], pr[2 * c + 1]);
            }
        }
        if (RECORD) {
            if (threadIdx.x == 0) {
                float* sc = slot + (size_t)(4 + T2_ROWS * S) * n;
                sc[0] = e2[0];
                sc[1] = e2[1];
                for (int q = 0; q < 4; ++q) sc[2 + q] = __int_as_float(pr[q]);
                if (D.mpair) {
                    for (int q = 0; q < 4; ++q) D.mpair[4 * t + q] = pr[q];
                }
            }
        }
        const float rsf = 1.0f / fminf(fmaxf(sqrtf(e2[0]), 0.01f), 2.0f);
        const float rss = 1.0f / fminf(fmaxf(sqrtf(e2[1]), 0.01f), 2.0f);
        const float Dgf = 1.0f / (th.sig_f * th.sig_f), Dgs = 1.0f / (th.sig_s * th.sig_s);
        const float sqf = kDiffAlpha / th.sig_f, sqs = kDiffAlpha / th.sig_s;
        float2* m2 = reinterpret_cast<float2*>(mbuf);
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) m2[row(r)] = make_float2(sqf * mf[r], sqs * ms[r]);
        for (int i = n + (int)threadIdx.x; i < p2; i += BS) m2[i] = make_float2(1.0e18f, 1.0e18f);
        const float sgn2 = C.paper_sign ? 2.0f : -2.0f;
        const float temp = C.temperature;
        float* it = slot + 4 * (size_t)n;
        // the recompute with a stored drift replays the positions: no records, no transport pass
        const bool replay = RECORD && !SELECT && D.drift != nullptr;
        for (int k = 0; k < S; ++k, it += T2_ROWS * (size_t)n) {
            const float beta = S > 1 ? 0.3f + 0.7f * (float)k / (float)(S - 1) : 1.0f;
            const float eff = 0.1f * sqrtf(beta);
            const float dstd = sqrtf(2.0f * eff * temp);
            __syncthreads();
            // mixture-prior sums (two components per float4 load)
            float hhf[R], hhs[R], off[R], sr[R], s1f[R], s1s[R], mff[R], mfs[R], mss[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                hhf[r] = sqf * hf[r];
                hhs[r] = sqs * hs[r];
                off[r] = 3.0e38f;
                sr[r] = s1f[r] = s1s[r] = mff[r] = mfs[r] = mss[r] = 0.f;
            }
            const int nq = (n + 1) >> 1;
            const float4* m4 = reinterpret_cast<const float4*>(mbuf);
            float* drift_k = D.drift ? D.drift + ((size_t)t * S + k) * kD2DriftRows * n : nullptr;
            if (replay) {
                // the forward's shift of the same positions (bit for bit what the pass below gives)
#pragma unroll
                for (int r = 0; r < R; ++r) off[r] = valid(r) ? drift_k[2 * (size_t)n + row(r)] : 0.0f;
            } else {
                for (int q = 0; q < nq; ++q) {
                    const float4 mm = m4[q];
#pragma unroll
                    for (int r = 0; r < R; ++r) {
                        const float a0 = hhf[r] - mm.x, b0 = hhs[r] - mm.y, a1 = hhf[r] - mm.z, b1 = hhs[r] - mm.w;
                        off[r] = fminf(off[r], fminf(fmaf(a0, a0, b0 * b0), fmaf(a1, a1, b1 * b1)));
                    }
                }
                if (SELECT && drift_k) {
#pragma unroll
                    for (int r = 0; r < R; ++r)
                        if (valid(r)) drift_k[2 * (size_t)n + row(r)] = off[r];
                }
            }
            for (int q = 0; q < nq; ++q) {
                const float4 mm = m4[q];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float df[2] = {hhf[r] - mm.x, hhf[r] - mm.z};
                    const float ds[2] = {hhs[r] - mm.y, hhs[r] - mm.w};
#pragma unroll
                    for (int c = 0; c < 2; ++c) {
                        const float e = __builtin_amdgcn_exp2f(fmaf(-df[c], df[c], fmaf(-ds[c], ds[c], off[r])));
                        const float ef = e * df[c], es = e * ds[c];
                        sr[r] += e;
                        s1f[r] += ef;
                        s1s[r] += es;
                        if (RECORD) {
                            mff[r] = fmaf(ef, df[c], mff[r]);
                            mfs[r] = fmaf(ef, ds[c], mfs[r]);
                            mss[r] = fmaf(es, ds[c], mss[r]);
                        }
                    }
                }
            }
            float uf[R], us[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                const float gpf = -((s1f[r] * (Dgf / sqf)) / (sr[r] + 1e-8f));
                const float gps = -((s1s[r] * (Dgs / sqs)) / (sr[r] + 1e-8f));
                const float hh = th.mu + hf[r] + hs[r];
                const float A = ysq / (safe_expf(hh) + 1e-8f) / nu;
                const float opa = 1.0f + A;
                const float gl = -0.5f + 0.5f * np1 * A / opa - C.lik_offset;
                const float gf = fminf(fmaxf(gpf + beta * gl, -10.0f), 10.0f);
                const float gs = fminf(fmaxf(gps + beta * gl, -10.0f), 10.0f);
                const float cl = 0.5f * np1 * A / (opa * opa);
                const float idet = 1.0f / (Dgf * Dgs + cl * (Dgf + Dgs));
                const float ca = (Dgs + cl) * idet, cb = -cl * idet, cd = (Dgf + cl) * idet;
                const float sf_ = 0.7f * (ca * gf + cb * gs), ss_ = 0.7f * (cb * gf + cd * gs);
                uf[r] = hf[r] * rsf;
                us[r] = hs[r] * rss;
                if (i < n) {
                    if (!replay) {
                        rec[i] = make_float4(uf[r], us[r], sf_, ss_);
                        rec2[i] = make_float4(sgn2 * ca * rsf, sgn2 * cb * rss, sgn2 * cb * rsf, sgn2 * cd * rss);
                    }
                    if (RECORD) {
                        const float isr = 1.0f / sr[r];
                        const float Ef = s1f[r] * isr, Es = s1s[r] * isr;
                        const float Mff = mff[r] * isr, Mss = mss[r] * isr;
                        it[T2_HF * (size_t)n + i] = hf[r];
                        it[T2_HS * (size_t)n + i] = hs[r];
                        it[T2_VF * (size_t)n + i] = vf[r];
                        it[T2_VS * (size_t)n + i] = vs[r];
                        it[T2_EF * (size_t)n + i] = Ef;
                        it[T2_ES * (size_t)n + i] = Es;
                        it[T2_VFF * (size_t)n + i] = Mff - Ef * Ef;
                        it[T2_VFS * (size_t)n + i] = mfs[r] * isr - Ef * Es;
                        it[T2_VSS * (size_t)n + i] = Mss - Es * Es;
                        it[T2_L * (size_t)n + i] = __log2f(sr[r]) - off[r];
                    }
                }
            }
            // transport
            float af[R], as[R];
#pragma unroll
            for (int r = 0; r < R; ++r) af[r] = as[r] = 0.f;
            if (!replay) __syncthreads();
            for (int j = 0; j < (replay ? 0 : n); ++j) {
                const float4 e = rec[j];
                const float4 q = rec2[j];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float df = uf[r] - e.x, ds = us[r] - e.y;
                    const float K = __builtin_amdgcn_rcpf(fmaf(df, df, fmaf(ds, ds, 1.0f)));
                    const float Pf = fmaf(q.x, df, q.y * ds), Ps = fmaf(q.z, df, q.w * ds);
                    af[r] = fmaf(K, fmaf(K, Pf, e.z), af[r]);
                    as[r] = fmaf(K, fmaf(K, Ps, e.w), as[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                float pf = af[r] * inv_n, ps = as[r] * inv_n;
                if (replay) {
                    pf = i < n ? drift_k[i] : 0.0f;
                    ps = i < n ? drift_k[n + i] : 0.0f;
                } else if (SELECT && drift_k && i < n) {
                    drift_k[i] = pf;
                    drift_k[n + i] = ps;
                }
                const float vnf = C.rms_rho * vf[r] + (1.0f - C.rms_rho) * pf * pf;
                const float vns = C.rms_rho * vs[r] + (1.0f - C.rms_rho) * ps * ps;
                float xa = 0.f, xb = 0.f;
                if (temp > 1e-6f && i < n) noise_xi(D, S, t, k, i, xa, xb);
                const float unf = hf[r] + eff * pf * rsqrtf(vnf + C.rms_eps) + dstd * xa;
                const float uns = hs[r] + eff * ps * rsqrtf(vns + C.rms_eps) + dstd * xb;
                if (RECORD && i < n) {
                    it[T2_PF * (size_t)n + i] = pf;
                    it[T2_PS * (size_t)n + i] = ps;
                    it[T2_UF * (size_t)n + i] = unf;
                    it[T2_US * (size_t)n + i] = uns;
                }
                vf[r] = vnf;
                vs[r] = vns;
                hf[r] = fminf(fmaxf(unf, -kD2Clamp), kD2Clamp);
                hs[r] = fminf(fmaxf(uns, -kD2Clamp), kD2Clamp);
            }
        }
        float sums[4] = {0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int r = 0; r < R; ++r) {
            xf[r] = hf[r];
            xs[r] = hs[r];
            if (valid(r)) {
                sums[0] += hf[r];
                sums[1] += hs[r];
                sums[2] += safe_expf(0.5f * (th.mu + hf[r] + hs[r]));
            }
        }
        block_sum<BS, 4>(sums, red);
        outv[SVPF_O2_PRED_LL] = pred_ll;
        outv[SVPF_O2_FMEAN] = sums[0] * inv_n;
        outv[SVPF_O2_SMEAN] = sums[1] * inv_n;
        outv[SVPF_O2_HMEAN] = th.mu + outv[SVPF_O2_FMEAN] + outv[SVPF_O2_SMEAN];
        outv[SVPF_O2_VOL] = sums[2] * inv_n;
        outv[SVPF_O2_BWF] = 1.0f / rsf;
        outv[SVPF_O2_BWS] = 1.0f / rss;
        outv[7] = 0.f;
        return pred_ll;
    }

    // score-side terms of one row at one Stein iteration, recomputed from the tape where needed: the
    // likelihood terms, the clamped score and its mask, and the 2x2 Newton inverse [[ca, cb], [cb, cd]]
    struct RowScore {
        float A, E, hsum, gf, gs, mf, ms, ca, cb, cd;
    };
    SVPF_DEV static RowScore row_score(const svpf_d2config& C, const D2Theta& th, float ysq, float Dgf, float Dgs,
                                       float sqf, float sqs, float beta, float hf, float hs, float Ef, float Es) {
        RowScore q;
        const float np1 = th.nu + 1.0f;
        q.hsum = th.mu + hf + hs;
        q.E = safe_expf(q.hsum);
        q.A = ysq / (q.E + 1e-8f) / th.nu;
        const float opa = 1.0f + q.A;
        const float gl = -0.5f + 0.5f * np1 * q.A / opa - C.lik_offset;
        const float grf = -Ef * (Dgf / sqf) + beta * gl, grs = -Es * (Dgs / sqs) + beta * gl;
        q.mf = (grf > -10.0f && grf < 10.0f) ? 1.0f : 0.0f;
        q.ms = (grs > -10.0f && grs < 10.0f) ? 1.0f : 0.0f;
        q.gf = fminf(fmaxf(grf, -10.0f), 10.0f);
        q.gs = fminf(fmaxf(grs, -10.0f), 10.0f);
        const float cl = 0.5f * np1 * q.A / (opa * opa);
        const float idet = 1.0f / (Dgf * Dgs + cl * (Dgf + Dgs));
        q.ca = (Dgs + cl) * idet;
        q.cb = -cl * idet;
        q.cd = (Dgf + cl) * idet;
        return q;
    }

    // ---- backward tick.  (xbf, xbs, vbf, vbs): adjoints after the tick, replaced by those before.
    // Per-row values that pass C does not read are recomputed from the tape after it instead of being
    // held in registers across it (one 1024-lane workgroup has 128 VGPRs a lane). ----
    SVPF_DEV void tick_back(const svpf_d2config& C, const D2Theta& th, const svpf_d2desc& D, int t, const float* slot,
                            float glbar, float (&xbf)[R], float (&xbs)[R], float (&vbf)[R], float (&vbs)[R],
                            double (&tb)[6]) {
        const int S = C.S;
        const float y_t = D.y[t];
        const float ysq = y_t * y_t;
        const float nu = th.nu, np1 = th.nu + 1.0f;
        const float* sc = slot + (size_t)(4 + T2_ROWS * S) * n;
        if (threadIdx.x < 6) scal[threadIdx.x] = sc[threadIdx.x];
        __syncthreads();
        const float e2f = scal[0], e2s = scal[1];
        const float rsf = 1.0f / fminf(fmaxf(sqrtf(e2f), 0.01f), 2.0f);
        const float rss = 1.0f / fminf(fmaxf(sqrtf(e2s), 0.01f), 2.0f);
        const float Dgf = 1.0f / (th.sig_f * th.sig_f), Dgs = 1.0f / (th.sig_s * th.sig_s);
        const float sqf = kDiffAlpha / th.sig_f, sqs = kDiffAlpha / th.sig_s;
        const float sgn2 = C.paper_sign ? 2.0f : -2.0f;
        const float bwf = 1.0f / rsf, bws = 1.0f / rss, isgf = 1.0f / th.sig_f, isgs = 1.0f / th.sig_s;

        float mbf[R], mbs[R], hbf[R], hbs[R], vbarf[R], vbars[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const bool ok = valid(r);
            mbf[r] = mbs[r] = 0.f;
            hbf[r] = ok ? xbf[r] : 0.f;
            hbs[r] = ok ? xbs[r] : 0.f;
            vbarf[r] = ok ? vbf[r] : 0.f;
            vbars[r] = ok ? vbs[r] : 0.f;
        }
        float rsbf = 0.f, rsbs = 0.f;
        float tmu = 0.f, tsf = 0.f, tss = 0.f, tnu = 0.f;
        for (int k = S - 1; k >= 0; --k) {
            const float* it = slot + 4 * (size_t)n + (size_t)k * T2_ROWS * n;
            const float beta = S > 1 ? 0.3f + 0.7f * (float)k / (float)(S - 1) : 1.0f;
            const float eff = 0.1f * sqrtf(beta);
            float uf[R], us[R], sfv[R], ssv[R], wf[R], ws[R], Qff[R], Qfs[R], Qsf[R], Qss[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                const bool ok = i < n;
                auto rd = [&](int row_) { return ok ? it[row_ * (size_t)n + i] : 0.f; };
                const float hf = rd(T2_HF), hs = rd(T2_HS), vvf = rd(T2_VF), vvs = rd(T2_VS);
                const float pf = rd(T2_PF), ps = rd(T2_PS), unf = rd(T2_UF), uns = rd(T2_US);
                const RowScore q = row_score(C, th, ysq, Dgf, Dgs, sqf, sqs, beta, hf, hs, rd(T2_EF), rd(T2_ES));
                sfv[r] = 0.7f * (q.ca * q.gf + q.cb * q.gs);
                ssv[r] = 0.7f * (q.cb * q.gf + q.cd * q.gs);
                Qff[r] = sgn2 * q.ca * rsf;
                Qfs[r] = sgn2 * q.cb * rss;
                Qsf[r] = sgn2 * q.cb * rsf;
                Qss[r] = sgn2 * q.cd * rss;
                // RMSProp + clamp, per coordinate
                {
                    const float vn = C.rms_rho * vvf + (1.0f - C.rms_rho) * pf * pf;
                    const float qq = rsqrtf(vn + C.rms_eps);
                    const float hm = (ok && unf > -kD2Clamp && unf < kD2Clamp) ? 1.0f : 0.0f;
                    const float hbb = hbf[r] * hm;
                    const float vt = vbarf[r] + hbb * (-0.5f * eff * pf * qq * qq * qq);
                    wf[r] = ok ? (hbb * eff * qq + vt * 2.0f * (1.0f - C.rms_rho) * pf) * inv_n : 0.f;
                    vbarf[r] = C.rms_rho * vt;
                    hbf[r] = hbb;
                }
                {
                    const float vn = C.rms_rho * vvs + (1.0f - C.rms_rho) * ps * ps;
                    const float qq = rsqrtf(vn + C.rms_eps);
                    const float hm = (ok && uns > -kD2Clamp && uns < kD2Clamp) ? 1.0f : 0.0f;
                    const float hbb = hbs[r] * hm;
                    const float vt = vbars[r] + hbb * (-0.5f * eff * ps * qq * qq * qq);
                    ws[r] = ok ? (hbb * eff * qq + vt * 2.0f * (1.0f - C.rms_rho) * ps) * inv_n : 0.f;
                    vbars[r] = C.rms_rho * vt;
                    hbs[r] = hbb;
                }
                uf[r] = hf * rsf;
                us[r] = hs * rss;
            }
            __syncthreads();  // the previous pass B finished reading rec / rec2
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!valid(r)) continue;
                const int i = row(r);
                rec[i] = make_float4(uf[r], us[r], -2.0f * sfv[r], -2.0f * ssv[r]);
                rec2[i] = make_float4(Qff[r], Qfs[r], Qsf[r], Qss[r]);
                rec3[i] = make_float4(wf[r], ws[r], 0.f, 0.f);
            }
            __syncthreads();
            // pass C
            float Gf[R], Gs[R], C1f[R], C1s[R], QBff[R], QBfs[R], QBsf[R], QBss[R], KWf[R], KWs[R], TCf[R], TCs[R];
#pragma unroll
            for (int r = 0; r < R; ++r)
                Gf[r] = Gs[r] = C1f[r] = C1s[r] = QBff[r] = QBfs[r] = QBsf[r] = QBss[r] = KWf[r] = KWs[r] = TCf[r] =
                    TCs[r] = 0.f;
            for (int j = 0; j < n; ++j) {
                const float4 e = rec[j];
                const float4 q = rec2[j];
                const float4 ww = rec3[j];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float df = uf[r] - e.x, ds = us[r] - e.y;
                    const float K = __builtin_amdgcn_rcpf(fmaf(df, df, fmaf(ds, ds, 1.0f)));
                    const float K2 = K * K;
                    // row part: w_i . df/dd with s_j (stored as -2 s_j), Qt_j:
                    //   K^2 (b + (-4 K ga - 2 al) d), b = Qt_j^T w_i, ga = b . d, al = w_i . s_j
                    const float al2 = fmaf(wf[r], e.z, ws[r] * e.w);
                    const float bf = fmaf(wf[r], q.x, ws[r] * q.z), bs = fmaf(wf[r], q.y, ws[r] * q.w);
                    const float ga = fmaf(bf, df, bs * ds);
                    const float co = fmaf(-4.0f * K, ga, al2);
                    Gf[r] = fmaf(K2, fmaf(co, df, bf), Gf[r]);
                    Gs[r] = fmaf(K2, fmaf(co, ds, bs), Gs[r]);
                    // column part with the weights w_j
                    C1f[r] = fmaf(K, ww.x, C1f[r]);
                    C1s[r] = fmaf(K, ww.y, C1s[r]);
                    const float k2f = K2 * ww.x, k2s = K2 * ww.y;
                    KWf[r] += k2f;
                    KWs[r] += k2s;
                    QBff[r] = fmaf(k2f, df, QBff[r]);
                    QBfs[r] = fmaf(k2f, ds, QBfs[r]);
                    QBsf[r] = fmaf(k2s, df, QBsf[r]);
                    QBss[r] = fmaf(k2s, ds, QBss[r]);
                    // sum_j K^3 w_j^T Qt_i d d: the third-order terms contracted with the row's own Qt
                    const float Pf = fmaf(Qff[r], df, Qfs[r] * ds), Ps = fmaf(Qsf[r], df, Qss[r] * ds);
                    const float k3p = K2 * K * fmaf(ww.x, Pf, ww.y * Ps);
                    TCf[r] = fmaf(k3p, df, TCf[r]);
                    TCs[r] = fmaf(k3p, ds, TCs[r]);
                }
            }
            float atf[R], ats[R], aaf[R], aas[R], bb[R], Lv[R], Efv[R], Esv[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                atf[r] = ats[r] = aaf[r] = aas[r] = bb[r] = Lv[r] = Efv[r] = Esv[r] = 0.f;
                if (!valid(r)) continue;  // rows past n see nonzero column sums: keep them out of the adjoints
                const int i = row(r);
                // column part of the position adjoint
                const float colf = -(2.0f * (sfv[r] * QBff[r] + ssv[r] * QBsf[r]) - 4.0f * TCf[r] +
                                     (KWf[r] * Qff[r] + KWs[r] * Qsf[r]));
                const float cols = -(2.0f * (sfv[r] * QBfs[r] + ssv[r] * QBss[r]) - 4.0f * TCs[r] +
                                     (KWf[r] * Qfs[r] + KWs[r] * Qss[r]));
                const float Uf = Gf[r] + colf, Us = Gs[r] + cols;
                hbf[r] += Uf * rsf;
                hbs[r] += Us * rss;
                rsbf += (uf[r] * Uf - (Qff[r] * QBff[r] + Qsf[r] * QBsf[r])) / rsf;
                rsbs += (us[r] * Us - (Qfs[r] * QBfs[r] + Qss[r] * QBss[r])) / rss;
                // the row's score terms again, from the tape
                const float Ef = it[T2_EF * (size_t)n + i], Es = it[T2_ES * (size_t)n + i];
                const RowScore q = row_score(C, th, ysq, Dgf, Dgs, sqf, sqs, beta, it[T2_HF * (size_t)n + i],
                                             it[T2_HS * (size_t)n + i], Ef, Es);
                // s = 0.7 Cinv g, Qt = sgn2 Cinv diag(rs)
                const float sbf = C1f[r], sbs = C1s[r];
                const float gbf = 0.7f * (q.ca * sbf + q.cb * sbs), gbs = 0.7f * (q.cb * sbf + q.cd * sbs);
                // Cinv adjoint (full 2x2): 0.7 sbar g^T + sgn2 Qbar, Qbar_ab = -QB_ab rs_b
                const float Xff = 0.7f * sbf * q.gf + sgn2 * (-QBff[r] * rsf);
                const float Xfs = 0.7f * sbf * q.gs + sgn2 * (-QBfs[r] * rss);
                const float Xsf = 0.7f * sbs * q.gf + sgn2 * (-QBsf[r] * rsf);
                const float Xss = 0.7f * sbs * q.gs + sgn2 * (-QBss[r] * rss);
                // Cbar = -Cinv X Cinv (Cinv = [[ca, cb], [cb, cd]])
                const float Yff = Xff * q.ca + Xfs * q.cb, Yfs = Xff * q.cb + Xfs * q.cd;
                const float Ysf = Xsf * q.ca + Xss * q.cb, Yss = Xsf * q.cb + Xss * q.cd;
                const float Cff = -(q.ca * Yff + q.cb * Ysf), Cfs = -(q.ca * Yfs + q.cb * Yss);
                const float Csf = -(q.cb * Yff + q.cd * Ysf), Css = -(q.cb * Yfs + q.cd * Yss);
                const float clb = Cff + Cfs + Csf + Css;
                tsf += Cff * (-2.0f * Dgf / th.sig_f);
                tss += Css * (-2.0f * Dgs / th.sig_s);
                const float gpbf = gbf * q.mf, gpbs = gbs * q.ms;
                const float glb = beta * (gpbf + gpbs);
                const float A = q.A, opa = 1.0f + A;
                const float Ab = glb * 0.5f * np1 / (opa * opa) + clb * 0.5f * np1 * (1.0f - A) / (opa * opa * opa);
                tnu += glb * 0.5f * A / opa + clb * 0.5f * A / (opa * opa) - Ab * A / nu;
                const float hbar_s = q.hsum < 20.0f ? Ab * (-A * q.E / (q.E + 1e-8f)) : 0.f;
                hbf[r] += hbar_s;
                hbs[r] += hbar_s;
                tmu += hbar_s;
                // mixture prior of the own particle
                const float Vff = it[T2_VFF * (size_t)n + i], Vfs = it[T2_VFS * (size_t)n + i];
                const float Vss = it[T2_VSS * (size_t)n + i];
                // d gp_c / d x_c' = -D_c (delta_cc' - V_cc' sq_c' / (sq_c alpha^2))
                const float Hf = -Dgf * gpbf * (1.0f - Vff * kDiffKappa) - Dgs * gpbs * (-Vfs * sqf / sqs * kDiffKappa);
                const float Hs = -Dgf * gpbf * (-Vfs * sqs / sqf * kDiffKappa) - Dgs * gpbs * (1.0f - Vss * kDiffKappa);
                hbf[r] += Hf;
                hbs[r] += Hs;
                // d gp / d sigma_c by homogeneity: scaling coordinate c of every particle, every mean and
                // sigma_c by lambda scales gp_c by 1/lambda and leaves the other coordinate's gp as it is, so
                //   sigma_c d(gpbar . gp)/dsigma_c = -gpbar_c gp_c - x_c xbar_c - sum_j m_jc mbar_jc
                // with xbar, mbar the prior's position and mean adjoints (pass B's sums); the means' part is
                // added once after the Stein loop, the means being fixed over it
                tsf -= (gpbf * (-Ef * (Dgf / sqf)) + uf[r] * bwf * Hf) * isgf;
                tss -= (gpbs * (-Es * (Dgs / sqs)) + us[r] * bws * Hs) * isgs;
                // pass B record
                aaf[r] = gpbf * Dgf;
                aas[r] = gpbs * Dgs;
                atf[r] = aaf[r] / sqf;
                ats[r] = aas[r] / sqs;
                bb[r] = atf[r] * Ef + ats[r] * Es;
                Lv[r] = it[T2_L * (size_t)n + i];
                Efv[r] = Ef;
                Esv[r] = Es;
            }
            __syncthreads();  // pass C finished reading rec / rec2 / rec3
#pragma unroll
            for (int r = 0; r < R; ++r) {
                if (!valid(r)) continue;
                const int i = row(r);
                rec[i] = make_float4(uf[r] / rsf * sqf, us[r] / rss * sqs, Lv[r], bb[r]);
                rec2[i] = make_float4(atf[r], ats[r], aaf[r], aas[r]);
            }
            // the scaled mixture means of the own rows (components), from the tick's start positions
            float mmf[R], mms[R];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int i = row(r);
                mmf[r] = i < n ? sqf * th.rho_f * slot[i] : 0.f;
                mms[r] = i < n ? sqs * th.rho_s * slot[n + i] : 0.f;
            }
            __syncthreads();
            // pass B: components = own rows; sum_j rt (a_j - psi kappa sq dd) with the constant
            // kappa sq applied once after the sum
            const float kf = sqf * kDiffKappa, ks = sqs * kDiffKappa;
            float accf[R], accs[R], apf[R], aps[R];
#pragma unroll
            for (int r = 0; r < R; ++r) accf[r] = accs[r] = apf[r] = aps[r] = 0.f;
            for (int j = 0; j < n; ++j) {
                const float4 e = rec[j];
                const float4 a = rec2[j];
#pragma unroll
                for (int r = 0; r < R; ++r) {
                    const float ddf = e.x - mmf[r], dds = e.y - mms[r];
                    const float rt = __builtin_amdgcn_exp2f(fmaf(-ddf, ddf, fmaf(-dds, dds, -e.z)));
                    const float rp = rt * fmaf(a.x, ddf, fmaf(a.y, dds, -e.w));
                    accf[r] = fmaf(rt, a.z, accf[r]);
                    accs[r] = fmaf(rt, a.w, accs[r]);
                    apf[r] = fmaf(rp, ddf, apf[r]);
                    aps[r] = fmaf(rp, dds, aps[r]);
                }
            }
#pragma unroll
            for (int r = 0; r < R; ++r) {
                mbf[r] += fmaf(-kf, apf[r], accf[r]);
                mbs[r] += fmaf(-ks, aps[r], accs[r]);
            }
        }
        // start positions and predicted particles of the tick, from the tape
        float xrf[R], xrs[R], uuf[R], uus[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            const bool ok = i < n;
            xrf[r] = ok ? slot[i] : 0.f;
            xrs[r] = ok ? slot[n + i] : 0.f;
            uuf[r] = ok ? slot[2 * (size_t)n + i] : 0.f;
            uus[r] = ok ? slot[3 * (size_t)n + i] : 0.f;
        }
        int pr[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) pr[q] = __float_as_int(scal[2 + q]);
        // bandwidths: ell_c^2 = (x_a - x_b)^2 of the predicted particles
        float hpf[R], hps[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hpf[r] = fminf(fmaxf(uuf[r], -kD2Clamp), kD2Clamp);
            hps[r] = fminf(fmaxf(uus[r], -kD2Clamp), kD2Clamp);
        }
        float rr[2] = {rsbf, rsbs};
        block_sum<BS, 2>(rr, red);
        for (int c = 0; c < 2; ++c) {
            const float e2 = c == 0 ? e2f : e2s;
            const int pa = pr[2 * c], pb = pr[2 * c + 1];
            const float sb = sqrtf(e2);
            if (pa < 0 || !(sb > 0.01f && sb < 2.0f)) continue;
            float* hl = reinterpret_cast<float*>(rec3);
            __syncthreads();
#pragma unroll
            for (int r = 0; r < R; ++r)
                if (valid(r)) hl[row(r)] = c == 0 ? hpf[r] : hps[r];
            __syncthreads();
            const float rsc = c == 0 ? rsf : rss;
            const float l2b = rr[c] * (-0.5f * rsc * rsc * rsc);
            const float dab = hl[pa] - hl[pb];
#pragma unroll
            for (int r = 0; r < R; ++r) {
                float& hb = c == 0 ? hbf[r] : hbs[r];
                if (row(r) == pa) hb += 2.0f * dab * l2b;
                if (row(r) == pb) hb -= 2.0f * dab * l2b;
            }
        }
        // predictive log-likelihood
        float lw[R], Ep[R], Ap[R], hh[R];
#pragma unroll
        for (int r = 0; r < R; ++r) {
            hh[r] = th.mu + hpf[r] + hps[r];
            Ep[r] = safe_expf(hh[r]);
            Ap[r] = ysq / (Ep[r] + 1e-8f) / nu;
            lw[r] = valid(r) ? th.st_const - 0.5f * hh[r] - np1 * 0.5f * log1pf(Ap[r]) : -3.0e38f;
        }
        float mx = -3.0e38f;
#pragma unroll
        for (int r = 0; r < R; ++r) mx = fmaxf(mx, lw[r]);
        mx = block_max<BS>(mx, red);
        float se[1] = {0.f};
#pragma unroll
        for (int r = 0; r < R; ++r)
            if (valid(r)) se[0] += expf(lw[r] - mx);
        block_sum<BS, 1>(se, red);
        const float inv_se = 1.0f / se[0];
        float trf = 0.f, trs = 0.f;
#pragma unroll
        for (int r = 0; r < R; ++r) {
            const int i = row(r);
            if (i >= n) continue;
            const float wt = expf(lw[r] - mx) * inv_se * glbar;
            const float opa = 1.0f + Ap[r];
            float dh = -0.5f;
            if (hh[r] < 20.0f) dh += 0.5f * np1 * Ap[r] / opa * (Ep[r] / (Ep[r] + 1e-8f));
            hbf[r] += wt * dh;
            hbs[r] += wt * dh;
            tmu += wt * dh;
            tnu += wt * (th.dst_const - 0.5f * log1pf(Ap[r]) + 0.5f * np1 * (Ap[r] / nu) / opa);
            // predict
            float zf, zs;
            noise_z(D, t, i, zf, zs);
            const float pmf = (uuf[r] > -kD2Clamp && uuf[r] < kD2Clamp) ? 1.0f : 0.0f;
            const float pms = (uus[r] > -kD2Clamp && uus[r] < kD2Clamp) ? 1.0f : 0.0f;
            // the means' part of the prior's sigma adjoint (homogeneity, above)
            tsf -= th.rho_f * xrf[r] * mbf[r] * isgf;
            tss -= th.rho_s * xrs[r] * mbs[r] * isgs;
            const float mtf = mbf[r] + hbf[r] * pmf, mts = mbs[r] + hbs[r] * pms;
            tsf += hbf[r] * pmf * zf;
            tss += hbs[r] * pms * zs;
            trf += mtf * xrf[r];
            trs += mts * xrs[r];
            xbf[r] = mtf * th.rho_f;
            xbs[r] = mts * th.rho_s;
            vbf[r] = vbarf[r];
            vbs[r] = vbars[r];
        }
        // per-lane partials of one tick in float, summed over ticks in double
        tb[0] += (double)tmu;
        tb[1] += (double)trf;
        tb[2] += (double)tsf;
        tb[3] += (double)trs;
        tb[4] += (double)tss;
        tb[5] += (double)tnu;
    }
};

}  // namespace svpf
