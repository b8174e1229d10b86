"""C ABI of the joint state-parameter filter without a GPU: the library exports
every function include/svpf_joint.h declares, the ctypes mirrors match the C
layouts, and the defaults are those of svpf_joint/src/svpf_joint.cu:21-72."""
import ctypes
import math
import os
import re
import subprocess

import pytest

from svpf_amd import _lib, joint

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "svpf_joint.h")


def test_library_exports_every_declared_function():
    src = re.sub(r"/\*.*?\*/", "", open(HEADER).read(), flags=re.S)
    names = {m.group(1) for m in re.finditer(r"^\s*[A-Za-z_][\w\s\*]*?\b(svpf_joint_\w+)\s*\(", src, flags=re.M)}
    assert len(names) >= 10
    assert names == set(joint.EXPORTED)
    L = joint.bind()
    for name in names:
        assert hasattr(L, name), name


def test_layouts_match_header(tmp_path):
    prog = tmp_path / "layout.c"
    lines = ["#include <stdio.h>", "#include <stddef.h>", '#include "svpf_joint.h"', "int main(void) {"]
    for cname, cls in (("SVPFJointConfig", joint.JointConfig), ("SVPFJointDiagnostics", joint.JointDiagnostics),
                       ("SVPFJointState", joint.JointState)):
        lines.append(f'printf("{cname}.size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname}.{f} %zu\\n", offsetof({cname}, {f}));')
    lines.append(f'printf("recw %d %d\\n", SVPF_JOINT_RECW, SVPF_JOINT_MAX_PARTICLES);')
    lines.append("return 0; }")
    prog.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-D__HIP_PLATFORM_AMD__", "-I", os.path.join(ROOT, "include"), "-I", "/opt/rocm/include",
                    str(prog), "-o", str(exe)], check=True)
    got = {}
    for line in subprocess.run([str(exe)], capture_output=True, text=True, check=True).stdout.split("\n"):
        if line.strip():
            k, *v = line.split()
            got[k] = [int(x) for x in v]
    for cname, cls in (("SVPFJointConfig", joint.JointConfig), ("SVPFJointDiagnostics", joint.JointDiagnostics),
                       ("SVPFJointState", joint.JointState)):
        assert got[f"{cname}.size"][0] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[f"{cname}.{f}"][0] == getattr(cls, f).offset, (cname, f)
    assert got["recw"] == [joint.RECW, joint.MAX_PARTICLES]
    assert len(joint.RECORD_FIELDS) == joint.RECW


def test_default_config_is_the_references():
    d = joint.default_config()
    assert (d["n_particles"], d["n_stein_steps"]) == (512, 5)
    want = dict(step_h=0.15, step_mu=0.005, step_rho=0.002, step_sigma=0.03, diffusion_mu=0.005,
                diffusion_rho=0.0005, diffusion_sigma=0.002, collapse_thresh_mu=0.05, collapse_thresh_rho=0.02,
                collapse_thresh_sigma=0.02, mu_prior_mean=-3.5, mu_prior_var=10.0, rho_prior_mean=2.0,
                rho_prior_var=5.0, sigma_prior_mean=-2.0, sigma_prior_var=5.0, nu=30.0, lik_offset=0.10,
                prior_weight=0.05)
    for k, v in want.items():
        assert d[k] == pytest.approx(v, rel=1e-7), k
    nu = 30.0
    assert d["student_t_const"] == pytest.approx(math.lgamma(15.5) - math.lgamma(15.0) - 0.5 * math.log(math.pi * nu),
                                                 abs=1e-5)  # float32 lgammaf difference
    # -E[log t_30^2] with psi(15) from its two-term expansion (svpf_joint.cu:58-64)
    psi15 = math.log(15.0) - 1.0 / 30.0 - 1.0 / (12.0 * 225.0)
    assert d["student_t_implied_offset"] == pytest.approx(-(math.log(nu) - 1.9635100260214235 - psi15), abs=2e-6)


@pytest.mark.parametrize("n", [0, -1, 4097])
def test_create_rejects_particle_counts_outside_the_range(n):
    with pytest.raises(ValueError):
        joint.JointFilter(n_particles=n)


def test_lds_bytes():
    L = joint.bind()
    assert L.svpf_joint_lds_bytes(4096) == 4096 * 32 + 512
    assert L.svpf_joint_lds_bytes(4096) <= 160 * 1024
    assert L.svpf_joint_lds_bytes(0) == -1 and L.svpf_joint_lds_bytes(4097) == -1
