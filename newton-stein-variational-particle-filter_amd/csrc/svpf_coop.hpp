// svpf_coop.hpp -- one SVPF filter spread over G cooperating workgroups.
//
// Latency form of the tick for a single filter: workgroup g owns rows
// [g*ROWS, (g+1)*ROWS) (ROWS = BS/CS, one row per CS adjacent lanes).  Work
// that is O(N) -- predict, guide, the median bandwidth, the mixture means --
// is done redundantly by every workgroup from identical inputs, so
// the only cross-workgroup traffic per tick is
//   * one exchange per Stein iteration of the transport records
//     {h/bw, s, c/bw} of every particle, and
//   * one exchange after the last iteration of the final positions plus each
//     workgroup's partial output sums.
// An exchange is a pair of global buffers (epoch parity) of 64-bit words
// holding {float value, uint32 epoch}: a word written with one 64-bit store
// carries its own validity, so readers need no barrier or fence -- they spin
// (bounded) until every word shows the current epoch.
#pragma once

#include "svpf_device.hpp"
#include "svpf_engine.hpp"

namespace svpf {

// Partial sums each workgroup publishes after the last iteration.
enum { PS_VOL = 0, PS_H = 1, PS_LWMAX = 2, PS_LWSUM = 3, PS_KSD = 4, PS_N = 8 };

struct CoopIO {
    const svpf_fparams* P;
    svpf_fstate* S;
    const svpf_fdesc* D;
    float* h;
    float* hp;
    float* v;
    uint64_t* xbuf;   // [2][4][n] transport records, [2][n] final positions, [2][G][PS_N] partials, then the
                      // commit area (svpf_coop_xbuf_words)
    int G;            // cooperating workgroups of the particle rows
    int g;            // this workgroup (g == G: the median workgroup, bwg launches only)
    int bwg;          // a median workgroup runs beside the G row workgroups
};

// Exchange buffer of one filter of n particles, in 64-bit words: the records,
// final positions and partials (G <= n / 8 + 2), then the commit stage -- each
// row's RMSProp state and log-weight as tagged words, 2n -- one commit counter,
// and [2] words of the median workgroup's pairwise-distance median.
__host__ __device__ inline size_t svpf_coop_stage_offset(int n) {
    return 10 * (size_t)n + 2 * (size_t)PS_N * ((size_t)n / 8 + 2);
}
__host__ __device__ inline size_t svpf_coop_med_offset(int n) { return svpf_coop_stage_offset(n) + 2 * (size_t)n + 1; }
__host__ __device__ inline size_t svpf_coop_xbuf_words(int n) { return svpf_coop_med_offset(n) + 2; }

SVPF_DEV uint64_t tag_word(float v, uint32_t epoch) { return ((uint64_t)epoch << 32) | (uint64_t)__float_as_uint(v); }

template <int BS, int CS>
struct CoopEngine {
    static constexpr int ROWS = BS / CS;
    static constexpr long kSpinLimit = 1L << 20;  // default bound of an exchange wait (polls)
    long spin_limit = kSpinLimit;
    uint32_t epoch_xor = 0;  // nonzero: the gathers wait for a tag nobody publishes (forced time-out)

    // LDS
    float4* hsc;     // [n]  transport records of all particles
    float* mbuf;     // [p2] scaled mixture means / median sort buffer
    float* hall;     // [n]  positions of all particles (predicted, then guided)
    float* hfin;     // [n]  final positions of the previous tick (this tick's h_prev)
    float* hpp;      // [n]  final positions two ticks back (MIM asymmetric rho)
    float* lwall;    // [n]  predictive log weights (all particles)
    float* red;      // [kRedWords]
    TickScalars* sc;
    int* flag;

    int n, p2, rg, cs, G, g, row0;
    float h, v, lw, s;  // own row

    __host__ __device__ static size_t lds_bytes(int nmax) {
        int p2 = LdsLayout::mbuf_len(nmax);
        size_t b = (size_t)nmax * 16 + (size_t)p2 * 4 + 4 * (size_t)nmax * 4 + kRedWords * 4 + sizeof(TickScalars) + 16;
        return (b + 15) & ~(size_t)15;
    }

    SVPF_DEV void bind(char* lds, int n_, int G_, int g_) {
        n = __builtin_amdgcn_readfirstlane(n_);
        p2 = LdsLayout::mbuf_len(n);
        G = G_;
        g = g_;
        hsc = reinterpret_cast<float4*>(lds);
        mbuf = reinterpret_cast<float*>(lds + (size_t)n * 16);
        hall = mbuf + p2;
        hfin = hall + n;
        hpp = hfin + n;
        lwall = hpp + n;
        red = lwall + n;
        sc = reinterpret_cast<TickScalars*>(red + kRedWords);
        flag = reinterpret_cast<int*>(sc + 1);
        rg = threadIdx.x / CS;
        cs = threadIdx.x % CS;
        row0 = g * ROWS;
    }
    SVPF_DEV int row() const { return row0 + rg; }
    SVPF_DEV bool valid() const { return row() < n; }

    // ---- exchange ----
    // Spin until words [0, cnt) of buf show `epoch`; store values through put(k, value).
    template <class Put>
    SVPF_DEV bool gather(const uint64_t* buf, int cnt, uint32_t epoch, Put put) {
        return gather_at(cnt, epoch, [buf](int k) { return buf + k; }, put);
    }
    // Words [0, cnt) at addresses addr(k).  Eight loads per thread are in flight at
    // once, so an exchange of up to 8 BS words (the records of n = 512 with full
    // Newton) costs one L2 round trip; each poll re-issues every word that is still
    // stale together, so a late producer costs one round trip per poll, not one per
    // word.
    template <class Addr, class Put>
    SVPF_DEV bool gather_at(int cnt, uint32_t epoch, Addr addr, Put put) {
        constexpr int U = 8;
        const uint32_t want = epoch ^ epoch_xor;
        bool ok = true;
        for (int k0 = threadIdx.x; k0 < cnt; k0 += U * BS) {
            uint64_t w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int k = k0 + u * BS;
                w[u] = k < cnt ? __hip_atomic_load(addr(k), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT)
                               : ((uint64_t)want << 32);
            }
            long spins = 0;
            for (;;) {
                bool stale = false;
#pragma unroll
                for (int u = 0; u < U; ++u) stale = stale || (uint32_t)(w[u] >> 32) != want;
                if (!stale) break;
                if (++spins > spin_limit) {
                    ok = false;
                    break;
                }
                __builtin_amdgcn_s_sleep(1);
#pragma unroll
                for (int u = 0; u < U; ++u)
                    if ((uint32_t)(w[u] >> 32) != want)  // k < cnt: padding words carry the epoch
                        w[u] = __hip_atomic_load(addr(k0 + u * BS), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            }
            if (!ok) break;
#pragma unroll
            for (int u = 0; u < U; ++u) {
                int k = k0 + u * BS;
                if (k < cnt) put(k, __uint_as_float((uint32_t)w[u]));
            }
        }
        if (!ok) *flag = 1;
        __syncthreads();
        bool all_ok = (*flag == 0);
        __syncthreads();
        return all_ok;
    }

    // The final exchange: n positions from xf into hfin, then G x PS_N partials from xp
    // into lwall (free by then).
    SVPF_DEV bool gather_final(const uint64_t* xf, const uint64_t* xp, uint32_t epoch) {
        float* hf = hfin;
        float* pr = lwall;
        const int nn = n;
        return gather_at(
            n + G * PS_N, epoch, [xf, xp, nn](int k) { return k < nn ? xf + k : xp + (k - nn); },
            [hf, pr, nn](int k, float val) {
                if (k < nn) hf[k] = val;
                else pr[k - nn] = val;
            });
    }

    SVPF_DEV void publish(uint64_t* p, float v, uint32_t epoch) {
        __hip_atomic_store(p, tag_word(v, epoch), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }

    // Reduction over all n entries of an LDS array (every workgroup redundantly).
    SVPF_DEV float lds_logmeanexp(const float* x) {
        float mx = -3.0e38f;
        for (int i = threadIdx.x; i < n; i += BS) mx = fmaxf(mx, x[i]);
        mx = block_max<BS>(mx, red);
        float se[1] = {0.f};
        for (int i = threadIdx.x; i < n; i += BS) se[0] += expf(x[i] - mx);
        block_sum<BS, 1>(se, red);
        return mx + logf(fmaxf(se[0] / (float)n, 1e-10f));
    }

    // Returns false when an exchange timed out (the filter's status is set).
    SVPF_DEV bool run(const CoopIO& io, int t0, int t1) {
        const svpf_fparams& P = *io.P;
        const svpf_fdesc& D = *io.D;
        const uint32_t F = P.flags;
        const bool rng_mode = (D.z_pred == nullptr);
        CounterRng rng{(uint32_t)D.seed, (uint32_t)(D.seed >> 32), D.filter_id};
        // coop_spin_limit > 0: bound of each exchange wait in polls; < 0: a forced time-out after
        // -coop_spin_limit polls (failure-path testing); 0: the default bound
        spin_limit = P.coop_spin_limit != 0 ? (long)abs(P.coop_spin_limit) : kSpinLimit;
        epoch_xor = P.coop_spin_limit < 0 ? 0x80000000u : 0u;
        __shared__ svpf_fstate fs;  // this workgroup's copy of the scalar state
        if (threadIdx.x == 0) {
            fs = *io.S;
            *flag = 0;
        }
        for (int i = threadIdx.x; i < n; i += BS) {
            hfin[i] = io.h[i];
            hpp[i] = io.hp[i];
        }
        v = valid() ? io.v[row()] : 0.f;
        lw = -3.0e38f;
        s = 0.f;
        __syncthreads();
        uint32_t epoch = fs.epoch + 1u;  // epochs are never reused across launches
        // the commit counter of this launch starts at 0 before workgroup 0 publishes anything
        // (the previous launch on the stream has finished with it)
        if (g == 0 && threadIdx.x == 0) {
            __hip_atomic_store(io.xbuf + svpf_coop_stage_offset(n) + 2 * (size_t)n, (uint64_t)0, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
            __builtin_amdgcn_s_waitcnt(0);  // the clear has completed before anything is published
        }
        uint64_t* xrec = io.xbuf;                          // [2][4][n]
        uint64_t* xfin = io.xbuf + 8 * (size_t)n;           // [2][n]
        uint64_t* xpart = xfin + 2 * (size_t)n;             // [2][G][PS_N]
        uint64_t* xmed = io.xbuf + svpf_coop_med_offset(n);  // [2]
        // bwg: the median workgroup (g == G) sorts the predicted particles and selects the
        // median pairwise distance while the row workgroups run the first gradient and
        // exchange its records unscaled; they scale them once the median arrives
        const bool bwg = io.bwg != 0 && P.bandwidth_mode == SVPF_BW_MEDIAN;
        const bool medwg = bwg && g == G;
        const float inv_n = 1.0f / (float)n;
        const int mode = (F & SVPF_F_NEWTON) ? ((F & SVPF_F_FULL_NEWTON) ? MODE_FULL : MODE_NEWTON) : MODE_PLAIN;
        const int nrec = (mode == MODE_FULL) ? 4 : 3;
        Engine<BS, 1, CS> E;  // pair loops of the single-workgroup engine, bound to this LDS
        E.hsc = hsc;
        E.mbuf = mbuf;
        E.n = n;
        E.p2 = p2;
        E.cs = cs;
        E.rg = rg;

        for (int t = t0; t < t1; ++t) {
            const uint32_t tick_id = (uint32_t)fs.timestep;
            SVPF_TRACE_POINT(t, 0);
            // ---- scalar prologue (identical in every workgroup) ----
            if (threadIdx.x == 0) prologue(P, fs, D.y[t]);
            __syncthreads();
            const float mu = sc->mu, sigma = sc->sigma, y_t = sc->y_t, y_prev = sc->y_prev;
            const float h_mean_prev = fs.h_mean_prev;

            // ---- predict every particle (src/svpf_opt_kernels.cu:125-276) ----
            {
                const bool mim = (F & (SVPF_F_MIM | SVPF_F_GUIDED)) != 0;
                const bool guided = (F & SVPF_F_GUIDED) != 0;
                const float rho_up = (F & SVPF_F_ASYM_RHO) ? P.rho_up : P.rho;
                const float rho_dn = (F & SVPF_F_ASYM_RHO) ? P.rho_down : P.rho;
                const float d_rho = (F & SVPF_F_LOCAL_PARAMS) ? P.delta_rho : 0.f;
                const float d_sig = (F & SVPF_F_LOCAL_PARAMS) ? P.delta_sigma : 0.f;
                float mean_implied = 0.f;
                if (guided) mean_implied = fmaxf(logf(y_t * y_t + 1e-10f) + P.implied_offset, -5.0f);
                for (int i = threadIdx.x; i < n; i += BS) {
                    float z, u = 1.0f;
                    if (rng_mode) {
                        uint32_t o[4];
                        rng.draw((uint32_t)i, tick_id, SVPF_PURPOSE_PREDICT, o);
                        z = normal_from_bits(o[0], o[1]);
                        u = u01_from_bits(o[2]);
                    } else {
                        z = D.z_pred[(size_t)t * n + i];
                        if (D.u_pred) u = D.u_pred[(size_t)t * n + i];
                    }
                    float hi = hfin[i], hpi = hpp[i];
                    float lev = P.gamma * y_prev / (safe_expf(hi * 0.5f) + 1e-8f);
                    float hn;
                    if (mim) {
                        float dev = hi - h_mean_prev;
                        float rho_adj = d_rho * tanhf(dev);
                        float sig_scale = 1.0f + d_sig * fabsf(dev);
                        float scale = (u < P.mim_p) ? P.mim_scale : 1.0f;
                        float br = (hi > hpi) ? rho_up : rho_dn;
                        float rr = fminf(fmaxf(br + rho_adj, 0.0f), 0.999f);
                        float mp = mu + rr * (hi - mu) + lev;
                        if (guided) {
                            float zs = (mean_implied - mp) / 2.5f;
                            float act = zs > P.ga_thr ? tanhf(zs - P.ga_thr) : 0.0f;
                            float a = P.ga_base + (P.ga_shock - P.ga_base) * act;
                            mp = (1.0f - a) * mp + a * mean_implied;
                        }
                        hn = clamp_logvol(mp + sigma * sig_scale * scale * z);
                    } else {
                        hn = clamp_logvol(mu + P.rho * (hi - mu) + sigma * z + lev);
                    }
                    float vol = safe_expf(hn);
                    float A = (y_t * y_t) / (vol + 1e-8f) / P.nu;
                    lwall[i] = P.st_const - 0.5f * hn - (P.nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f));
                    if (F & SVPF_F_GUIDE) {
                        const float gs = sc->gs, gm = sc->guide_mean;
                        if (F & SVPF_F_GUIDE_PRESERVING) {
                            float nm = (1.0f - gs) * h_mean_prev + gs * gm;
                            hn = clamp_logvol(nm + (hn - h_mean_prev));
                        } else {
                            hn = clamp_logvol(hn + gs * (gm - hn));
                        }
                    }
                    hall[i] = hn;
                    if (g == 0 && D.h_pred) D.h_pred[i] = hn;
                }
            }
            __syncthreads();
            SVPF_TRACE_POINT(t, 1);
            const uint32_t ep_med = epoch;  // the tick's median word carries the first records epoch
            if (medwg) {
                // ---- the median workgroup: the bandwidth, then the row workgroups' final exchange ----
                SVPF_TRACE_STAMP(t, 44);
                bandwidth(P, fs, y_t, t);
                if (threadIdx.x == 0) publish(xmed + (ep_med & 1u), fs.med_r2, ep_med);
                SVPF_TRACE_STAMP(t, 45);
                // its scalar state follows the row workgroups' exactly, so it can commit the launch
                const float pred_ll_m = lds_logmeanexp(lwall);
                epoch += (uint32_t)sc->budget;  // one records exchange per Stein iteration of the rows
                const uint32_t ep = epoch++;
                for (int i = threadIdx.x; i < n; i += BS) hpp[i] = hfin[i];
                __syncthreads();
                if (!gather_final(xfin + (size_t)(ep & 1u) * n, xpart + (size_t)(ep & 1u) * G * PS_N, ep))
                    return false;
                if (threadIdx.x < kWave) finish_from_partials(P, fs, D, t, pred_ll_m, sc->budget);
                __syncthreads();
                continue;
            }
            const float pred_ll = lds_logmeanexp(lwall);
            h = valid() ? hall[row()] : 0.f;
            SVPF_TRACE_POINT(t, 2);

            // ---- bandwidth (every workgroup, all particles; with bwg the median workgroup's) ----
            bool bw_ready = !bwg;
            if (!bwg) bandwidth(P, fs, y_t, t);
            float ib = sc->ib, rs = sc->rs;  // with bwg: set when the median arrives
            SVPF_TRACE_POINT(t, 3);

            // ---- scaled mixture means (unsorted: at this size a sort per tick costs more than
            //      the nearest-component pass it would replace) ----
            const float inv_s2 = 1.0f / (sigma * sigma);
            const float sq = sqrtf(1.4426950408889634f * 0.5f * inv_s2);
            {
                const float lev_num = P.gamma * y_prev;
                for (int i = threadIdx.x; i < n; i += BS)
                    mbuf[i] = sq * (mu + P.rho * (hfin[i] - mu) + lev_num / (expf(hfin[i] * 0.5f) + 1e-8f));
                for (int i = n + threadIdx.x; i < p2; i += BS) mbuf[i] = 1.0e18f;
                __syncthreads();
            }
            SVPF_TRACE_POINT(t, 4);

            // ---- annealed Stein loop ----
            const int n_anneal = (F & SVPF_F_ANNEAL) ? max(P.n_anneal, 1) : 1;
            const int budget = sc->budget;
            const float temp = (F & SVPF_F_SVLD) ? P.temperature : 0.0f;
            const float sgn2 = (F & SVPF_F_PAPER_SIGN) ? 2.0f : -2.0f;
            int total = 0;
            float ksdp = 0.f;
            for (int ai = 0; ai < n_anneal; ++ai) {
                float beta;
                if (!(F & SVPF_F_ANNEAL)) beta = 1.0f;
                else if (F & SVPF_F_LINEAR_BETA) beta = n_anneal > 1 ? 0.3f + 0.7f * (float)ai / (float)(n_anneal - 1) : 1.0f;
                else if (F & SVPF_F_ADAPT_BETA) beta = Engine<BS, 1, CS>::adaptive_beta(fs.ksd_prev, ai, n_anneal);
                else beta = (ai % 3 == 0) ? 0.3f : ((ai % 3 == 1) ? 0.65f : 1.0f);
                const float eff = P.base_step * sqrtf(beta);
                const float dstd = sqrtf(2.0f * eff * temp);
                int si = budget / n_anneal;
                if (ai == n_anneal - 1) si = budget - si * (n_anneal - 1);
                for (int s_it = 0; s_it < si; ++s_it) {
                    const bool last = (ai == n_anneal - 1) && (s_it == si - 1);
                    const int it_row = total;
                    ++total;
                    __syncthreads();  // the previous transport finished reading hsc
                    // gradient of the own row (src/svpf_opt_kernels.cu:354-521)
                    float sr[1], srd[1];
                    E.h[0] = h;
                    const float own_mm[1] = {valid() ? mbuf[row()] : h * sq};  // the row's own component
                    E.template prior_sums<false, true>(sq, sr, srd, own_mm);
                    float sj, cj, cv;
                    row_gradient(P, beta, y_t, inv_s2, sq, sgn2, mode, sr[0], srd[0], sj, cj, cv);
                    SVPF_TRACE_POINT(t, 8 + 3 * it_row);
                    const uint32_t ep = epoch++;
                    uint64_t* xr = xrec + (size_t)(ep & 1u) * 4 * n;
                    if (cs == 0 && valid()) {
                        int i = row();
                        publish(xr + i, bw_ready ? h * rs : h, ep);
                        publish(xr + n + i, sj, ep);
                        publish(xr + 2 * (size_t)n + i, bw_ready ? cj * rs : cj, ep);
                        if (nrec == 4) publish(xr + 3 * (size_t)n + i, cv, ep);
                    }
                    float4* rec = hsc;
                    const int nn = n;
                    if (!gather(xr, nrec * n, ep, [rec, nn](int k, float val) {
                            int a = k / nn, i = k - a * nn;
                            reinterpret_cast<float*>(&rec[i])[a] = val;
                        }))
                        return false;
                    if (!bw_ready) {
                        // the records went out unscaled: scale them once the median is in
                        // (one rounding each, as when scaled before publishing)
                        if (!await_bandwidth(xmed, ep_med, fs)) return false;
                        ib = sc->ib;
                        rs = sc->rs;
                        for (int i = threadIdx.x; i < n; i += BS) {
                            hsc[i].x *= rs;
                            hsc[i].z *= rs;
                        }
                        __syncthreads();
                        bw_ready = true;
                    }
                    SVPF_TRACE_POINT(t, 9 + 3 * it_row);
                    // transport of the own row
                    float phi, kp;
                    if (mode == MODE_NEWTON) {
                        if (last) row_phi<MODE_NEWTON, true>(E, ib, rs, inv_n, phi, kp);
                        else row_phi<MODE_NEWTON, false>(E, ib, rs, inv_n, phi, kp);
                    } else if (mode == MODE_FULL) {
                        if (last) row_phi<MODE_FULL, true>(E, ib, rs, inv_n, phi, kp);
                        else row_phi<MODE_FULL, false>(E, ib, rs, inv_n, phi, kp);
                    } else {
                        if (last) row_phi<MODE_PLAIN, true>(E, ib, rs, inv_n, phi, kp);
                        else row_phi<MODE_PLAIN, false>(E, ib, rs, inv_n, phi, kp);
                    }
                    if (last) ksdp = kp;
                    float vn = P.rms_rho * v + (1.0f - P.rms_rho) * phi * phi;
                    float drift = eff * phi * rsqrtf(vn + P.rms_eps);
                    float diff = 0.0f;
                    if (temp > 1e-6f && valid()) {
                        int i = row();
                        float xi;
                        if (rng_mode) xi = rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_SVLD0 + (uint32_t)it_row);
                        else xi = (it_row < D.xi_rows) ? D.xi[((size_t)t * D.xi_rows + it_row) * n + i] : 0.0f;
                        diff = dstd * xi;
                    }
                    v = vn;
                    h = clamp_logvol(h + drift + diff);
                    SVPF_TRACE_POINT(t, 10 + 3 * it_row);
                }
            }

            if (!bw_ready && !await_bandwidth(xmed, ep_med, fs)) return false;  // no Stein iteration ran

            // ---- rejuvenation (src/svpf_opt_kernels.cu:1248-1275) ----
            if ((F & SVPF_F_REJUV) && (F & SVPF_F_GUIDE) && fs.timestep > 10 && fs.ksd_prev > P.rej_thr && valid()) {
                int i = row();
                float u = D.u_rej ? D.u_rej[(size_t)t * n + i] : rng.uniform((uint32_t)i, tick_id, SVPF_PURPOSE_REJUV);
                if (u < P.rej_p) {
                    float z = D.z_rej ? D.z_rej[(size_t)t * n + i] : rng.normal((uint32_t)i, tick_id, SVPF_PURPOSE_REJUV);
                    h = clamp_logvol((1.0f - P.rej_blend) * h + P.rej_blend * (sc->guide_mean + sc->guide_std * z));
                }
            }

            // ---- final exchange: positions + partial output sums ----
            if (total == 0) lw = valid() ? lwall[row()] : -3.0e38f;
            float lmax = (valid() && cs == 0) ? lw : -3.0e38f;
            lmax = block_max<BS>(lmax, red);
            float part[4] = {0.f, 0.f, 0.f, 0.f};
            if (valid() && cs == 0) {
                part[0] = safe_expf(h * 0.5f);
                part[1] = h;
                part[2] = expf(lw - lmax);
                part[3] = ksdp;
            }
            block_sum<BS, 4>(part, red);
            SVPF_TRACE_POINT(t, 40);
            const uint32_t ep = epoch++;
            uint64_t* xf = xfin + (size_t)(ep & 1u) * n;
            uint64_t* xp = xpart + (size_t)(ep & 1u) * G * PS_N;
            if (cs == 0 && valid()) publish(xf + row(), h, ep);
            if (threadIdx.x < PS_N) {  // every one of the PS_N slots is published (the gather waits on all)
                const int k = threadIdx.x;
                float val = 0.f;
                if (k == PS_VOL) val = part[0];
                else if (k == PS_H) val = part[1];
                else if (k == PS_LWMAX) val = lmax;
                else if (k == PS_LWSUM) val = part[2];
                else if (k == PS_KSD) val = part[3];
                publish(xp + g * PS_N + k, val, ep);
            }
            // the previous final positions become this tick's h_prev
            for (int i = threadIdx.x; i < n; i += BS) hpp[i] = hfin[i];
            __syncthreads();
            // failure-path testing: one workgroup's waits of the launch's final exchange give
            // up after 4 polls, after it has published its own rows
            if (g == P.coop_fault_wg && t == t1 - 1) {
                epoch_xor = 0x80000000u;
                spin_limit = 4;
            }
            // final positions and partials in one gather (one round of polls)
            if (!gather_final(xf, xp, ep)) return false;
            SVPF_TRACE_POINT(t, 41);
            if (threadIdx.x < kWave) finish_from_partials(P, fs, D, t, pred_ll, total);
            __syncthreads();
            SVPF_TRACE_POINT(t, 42);
        }

        // ---- commit: all or nothing.  Each workgroup stages its rows' RMSProp state and
        //      log-weight as tagged words and arrives at the commit counter; the one
        //      arriving last (every other workgroup has passed every exchange) gathers the
        //      staged rows -- tagged, so no fence orders them -- and writes the filter's
        //      particles, which every workgroup holds after the final exchange, those rows
        //      and the scalars.  A launch in which any wait timed out never completes the
        //      count, or fails the last gather, and leaves the state as it was. ----
        uint64_t* stage = io.xbuf + svpf_coop_stage_offset(n);
        uint64_t* commit = stage + 2 * (size_t)n;
        const uint32_t ep_c = epoch;  // one past the launch's last exchange
        __shared__ int last_wg;
        if (cs == 0 && valid()) {
            publish(stage + row(), v, ep_c);
            publish(stage + n + row(), lw, ep_c);
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            // workgroup 0 cleared the counter before its first publication, which every
            // workgroup arriving here has gathered: one atomic add each, no retries
            const uint64_t before = __hip_atomic_fetch_add(commit, (uint64_t)1, __ATOMIC_RELAXED,
                                                           __HIP_MEMORY_SCOPE_AGENT);
            // the median workgroup arrives too (it has no rows to stage): a launch whose median
            // workgroup timed out commits nothing either
            last_wg = (before == (uint64_t)(G + (bwg ? 1 : 0) - 1)) ? 1 : 0;
        }
        __syncthreads();
        if (last_wg) {
            float* staged = reinterpret_cast<float*>(hsc);  // [2][n]; the records are dead now
            if (!gather(stage, 2 * n, ep_c, [staged](int k, float val) { staged[k] = val; })) return false;
            for (int i = threadIdx.x; i < n; i += BS) {
                io.h[i] = hfin[i];
                io.hp[i] = hpp[i];
                io.v[i] = staged[i];
                if (D.log_w) D.log_w[i] = staged[n + i];
            }
            if (threadIdx.x == 0) {
                fs.epoch = epoch;
                *io.S = fs;
            }
        }
        return true;
    }

    // The median workgroup's pairwise-distance median of this tick (epoch ep): the
    // bandwidth and its state, as bandwidth() sets them.  All threads call it.
    SVPF_DEV bool await_bandwidth(const uint64_t* xmed, uint32_t ep, svpf_fstate& S) {
        TickScalars* t = sc;
        if (!gather(xmed + (ep & 1u), 1, ep, [t](int, float val) { t->bw_sq = val; })) return false;
        if (threadIdx.x == 0) {
            const float med = sc->bw_sq;
            S.med_r2 = med;
            const float bw = fminf(fmaxf(sqrtf(med), 0.01f), 2.0f);
            sc->bw = bw;
            sc->ib = 1.0f / (bw * bw);
            sc->rs = 1.0f / bw;
        }
        __syncthreads();
        return true;
    }

    // Outputs of the tick from the G row workgroups' partials (in lwall) and the host
    // tail.  Called by the 64 lanes of wave 0: lane l reduces workgroups l, l + 64, ...
    SVPF_DEV void finish_from_partials(const svpf_fparams& P, svpf_fstate& S, const svpf_fdesc& D, int t,
                                       float pred_ll, int total) {
        const float* pr = lwall;
        const float inv_n = 1.0f / (float)n;
        const int l = threadIdx.x;
        float M = -3.0e38f;
        for (int q = l; q < G; q += kWave) M = fmaxf(M, pr[q * PS_N + PS_LWMAX]);
        M = wave_max(M);
        float vs = 0.f, hs = 0.f, ks = 0.f, ls = 0.f;
        for (int q = l; q < G; q += kWave) {
            vs += pr[q * PS_N + PS_VOL];
            hs += pr[q * PS_N + PS_H];
            ks += pr[q * PS_N + PS_KSD];
            float mq = pr[q * PS_N + PS_LWMAX];
            if (mq > -1.0e38f) ls += pr[q * PS_N + PS_LWSUM] * expf(mq - M);
        }
        vs = wave_sum(vs);
        hs = wave_sum(hs);
        ks = wave_sum(ks);
        ls = wave_sum(ls);
        if (l == 0) {
            float ll = M + logf(fmaxf(ls * inv_n, 1e-10f));
            float vol = vs * inv_n, hm = hs * inv_n;
            float ksd = sqrtf(fmaxf(ks * inv_n * inv_n, 0.0f));
            finish(P, S, D, t, ll, vol, hm, ksd, pred_ll, total);
        }
    }

    // Row gradient (src/svpf_opt_kernels.cu:354-521); sets lw and s of the own row.
    SVPF_DEV void row_gradient(const svpf_fparams& P, float beta, float y_t, float inv_s2, float sq, float sgn2,
                               int mode, float sr, float srd, float& sj, float& cj, float& cv) {
        const uint32_t F = P.flags;
        const float nu = P.nu;
        const float eb = (F & SVPF_F_FAN) ? 1.0f : beta;
        const float ysq = y_t * y_t;
        float gp = -((srd * (inv_s2 / sq)) / (sr + 1e-8f));
        float vol = safe_expf(h);
        float A = ysq / (vol + 1e-8f) / nu;
        float opa = 1.0f + A;
        lw = (F & SVPF_F_FAN) ? 0.0f : P.st_const - 0.5f * h - (nu + 1.0f) * 0.5f * log1pf(fmaxf(A, -0.999f));
        if (!valid()) lw = -3.0e38f;
        float gl;
        if (F & SVPF_F_EXACT_GRAD) gl = -0.5f + 0.5f * (nu + 1.0f) * A / opa - P.lik_offset;
        else gl = (logf(ysq + 1e-10f) - h + P.lik_offset) / 1.4f;
        float gg = fminf(fmaxf(gp + eb * gl, -10.0f), 10.0f);
        sj = gg;
        cj = sgn2;
        cv = 1.0f;
        if (mode != MODE_PLAIN) {
            float hl = -0.5f * (nu + 1.0f) * A / (opa * opa);
            cv = fminf(fmaxf(-(hl - inv_s2), 0.1f), 100.0f);
            if (mode == MODE_NEWTON) {
                sj = 0.7f * gg / cv;
                cj = sgn2 / cv;
            }
        }
        s = sj;
    }

    template <int MODE, bool KSD>
    SVPF_DEV void row_phi(Engine<BS, 1, CS>& E, float ib, float rs, float inv_n, float& phi, float& kp) {
        float acc[1], a1[1], a2[1], a3[1], a4[1], hw[1], kn[1];
        E.h[0] = h;
        E.template transport_sums<MODE, KSD>(ib, rs, acc, a1, a2, a3, a4, hw, kn);
        float p = acc[0] * inv_n;
        if (MODE == MODE_FULL) {
            float hb = fminf(fmaxf(hw[0] / fmaxf(kn[0], 1e-6f), 0.1f), 100.0f);
            p = p / hb * 0.7f;
        }
        phi = p;
        kp = KSD ? (s * a1[0] + 2.0f * rs * (s * a2[0] - a3[0]) + 2.0f * ib * a4[0]) : 0.f;
    }

    // Scalar prologue (src/svpf_optimized_graph.cu:932-1138) on a local state copy.
    SVPF_DEV void prologue(const svpf_fparams& P, svpf_fstate& S, float y_t) {
        const uint32_t F = P.flags;
        float y_prev = S.y_prev;
        float mu = (F & SVPF_F_ADAPT_MU) ? S.mu_state : P.mu;
        float sigma = P.sigma_z;
        if ((F & SVPF_F_ADAPT_SIGMA) && S.timestep > 0) {
            float ve = fmaxf(S.vol_prev, 1e-4f);
            float rz = fabsf(y_t) / ve;
            float boost = 1.0f;
            if (rz > P.sb_thr) boost = 1.0f + (P.sb_max - 1.0f) * fminf((rz - P.sb_thr) / 3.0f, 1.0f);
            sigma = P.sigma_z * boost;
            S.sigma_z_eff = sigma;
        }
        float gs = P.gs_base;
        if (F & SVPF_F_GUIDE) {
            if ((F & SVPF_F_ADAPT_GUIDE) && S.timestep > 0) {
                float ve = fmaxf(S.vol_prev, 1e-4f);
                float rz = fabsf(y_t) / ve;
                float ih = logf(y_t * y_t + 1e-8f) + P.implied_offset;
                float he = logf(ve * ve + 1e-8f);
                if (ih - he > 0.0f && rz > P.gs_thr)
                    gs = P.gs_base + (P.gs_max - P.gs_base) * fminf((rz - P.gs_thr) / 3.0f, 1.0f);
            }
            if (!S.guide_initialized) {
                S.guide_mean = mu;
                S.guide_var = P.sigma_z * P.sigma_z / (1.0f - P.rho * P.rho);
                S.guide_initialized = 1;
            }
            float m_pred = mu + P.rho * (S.guide_mean - mu);
            float P_pred = P.rho * P.rho * S.guide_var + P.sigma_z * P.sigma_z;
            float log_y2 = logf(y_t * y_t + 1e-8f);
            float Sv = P_pred + (4.93f + 2.0f);
            float Kg = P_pred / (Sv + 1e-8f);
            float innov = log_y2 - (m_pred - P.implied_offset);
            S.guide_mean = m_pred + Kg * innov;
            S.guide_var = (1.0f - Kg) * P_pred;
        }
        int budget;
        if (S.timestep < 10) budget = P.stein_max;
        else {
            float x = fminf(fmaxf((S.ksd_prev - 0.05f) / (0.50f - 0.05f), 0.0f), 1.0f);
            budget = P.stein_min + (int)(x * (float)(P.stein_max - P.stein_min));
        }
        sc->mu = mu;
        sc->sigma = sigma;
        sc->y_t = y_t;
        sc->y_prev = y_prev;
        sc->gs = gs;
        sc->guide_mean = S.guide_mean;
        sc->guide_std = sqrtf(fmaxf(S.guide_var, 1e-6f));
        sc->budget = budget;
    }

    // Outputs and host tail (src/svpf_opt_kernels.cu:1109-1153, src/svpf_optimized_graph.cu:1383-1388).
    SVPF_DEV void finish(const svpf_fparams& P, svpf_fstate& S, const svpf_fdesc& D, int t, float ll, float vol,
                         float hm, float ksd, float pred_ll, int total) {
        if (g == 0 && D.out) {
            float* o = D.out + (size_t)t * SVPF_OUTW;
            o[SVPF_O_LOGLIK] = ll;
            o[SVPF_O_VOL] = vol;
            o[SVPF_O_HMEAN] = hm;
            o[SVPF_O_KSD] = ksd;
            o[SVPF_O_BW] = sc->bw;
            o[SVPF_O_PRED_LL] = pred_ll;
            o[SVPF_O_STEPS] = (float)total;
            o[SVPF_O_BWSQ] = sc->bw_sq;
        }
        S.last_loglik = ll;
        S.last_vol = vol;
        S.last_h_mean = hm;
        S.last_bw = sc->bw;
        S.last_ksd = ksd;
        S.last_pred_loglik = pred_ll;
        S.last_guide_strength = sc->gs;
        S.stein_steps_used = total;
        S.h_mean_prev = hm;
        S.vol_prev = vol;
        S.ksd_prev = ksd;
        if ((P.flags & SVPF_F_ADAPT_MU) && S.timestep > 10) {
            float Pp = S.mu_var + P.mu_q;
            float Rm = P.mu_rscale * sc->bw * sc->bw;
            float Kg = Pp / (Pp + Rm + 1e-8f);
            float mn = S.mu_state + Kg * (hm - S.mu_state);
            S.mu_state = fminf(fmaxf(mn, P.mu_min), P.mu_max);
            S.mu_var = (1.0f - Kg) * Pp;
        }
        S.y_prev = sc->y_t;
        S.timestep += 1;
    }

    // Bandwidth over all particles in hall (every workgroup).
    SVPF_DEV void bandwidth(const svpf_fparams& P, svpf_fstate& S, float y_t, int t) {
        if (P.bandwidth_mode == SVPF_BW_MEDIAN) {
            // the transport records are not live yet this tick: their LDS is the
            // sort input, then the row brackets and the candidate buffer
            float* tmp = reinterpret_cast<float*>(hsc);
            for (int i = threadIdx.x; i < n; i += BS) tmp[i] = hall[i];
            __syncthreads();
            block_sort<BS>(tmp, mbuf, n);
            SVPF_TRACE_POINT(t, 50);
            int* plo = reinterpret_cast<int*>(hsc);
            float med = median_pairwise_r2<BS, 2>(mbuf, n, S.med_r2, plo, plo + n, reinterpret_cast<float*>(plo + 2 * n),
                                                  2 * n, reinterpret_cast<uint32_t*>(red));
            if (threadIdx.x == 0) {
                S.med_r2 = med;
                float bw = fminf(fmaxf(sqrtf(med), 0.01f), 2.0f);
                sc->bw = bw;
                sc->ib = 1.0f / (bw * bw);
                sc->rs = 1.0f / bw;
                sc->bw_sq = med;
            }
            __syncthreads();
            return;
        }
        float sums[2] = {0.f, 0.f};
        float mn = 3.0e38f, mx = -3.0e38f;
        for (int i = threadIdx.x; i < n; i += BS) {
            float x = hall[i];
            sums[0] += x;
            sums[1] += x * x;
            mn = fminf(mn, x);
            mx = fmaxf(mx, x);
        }
        block_sum<BS, 2>(sums, red);
        mn = block_min<BS>(mn, red);
        mx = block_max<BS>(mx, red);
        if (threadIdx.x == 0) {
            float inv_n = 1.0f / (float)n;
            float mean = sums[0] * inv_n;
            float var = sums[1] * inv_n - mean * mean;
            float bw;
            if (P.bandwidth_mode == SVPF_BW_SILVERMAN) {
                float sd = sqrtf(fmaxf(var, 1e-6f));
                bw = fminf(fmaxf(1.06f * sd * powf((float)n, -0.2f), 0.05f), 2.0f);
                sc->bw_sq = bw * bw;
            } else {
                float spread = mx - mn;
                float bsn = fmaxf(2.0f * var / logf((float)n + 1.0f), 1e-6f);
                float prev = S.bw_sq_ema;
                float bsq = prev > 0.0f ? 0.3f * bsn + 0.7f * prev : bsn;
                float ar = fabsf(y_t);
                float re = S.return_ema, rv = S.return_var;
                re = re > 0.0f ? 0.05f * ar + 0.95f * re : ar;
                rv = rv > 0.0f ? 0.05f * ar * ar + 0.95f * rv : ar * ar;
                S.return_ema = re;
                S.return_var = rv;
                float vr = ar / fmaxf(re, 1e-8f);
                float sf = fminf(spread * 0.5f, 2.0f);
                float comb = fmaxf(vr, sf);
                float scale = fmaxf(fminf(1.0f - 0.25f * fminf(comb - 1.0f, 2.0f), 1.0f), 0.5f);
                bsq *= scale;
                S.bw_sq_ema = bsq;
                bw = fminf(fmaxf(sqrtf(bsq), 0.01f), 2.0f);
                sc->bw_sq = bsq;
            }
            sc->bw = bw;
            sc->ib = 1.0f / (bw * bw);
            sc->rs = 1.0f / bw;
        }
        __syncthreads();
    }
};

}  // namespace svpf
