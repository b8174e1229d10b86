"""The reference's Python module name (py-torch/svpf_python.cpp:483, PYBIND11_MODULE(pysvpf, m)).

With this directory on sys.path (as for svpf_amd), ``import pysvpf`` gives the
same names the reference's extension exports -- SVPF, SVPFResult -- backed by
libsvpf_amd.so through svpf_amd.pysvpf.
"""
from svpf_amd.pysvpf import *  # noqa: F401,F403
from svpf_amd.pysvpf import __all__  # noqa: F401
