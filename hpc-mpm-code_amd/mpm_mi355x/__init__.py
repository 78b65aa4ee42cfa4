"""Python host bindings of the MI355X explicit-MPM step engine (C ABI in
include/mpm/mpm_capi.h, library hpc-mpm-code_amd/libmpm_mi355x.so)."""
from . import capi, scenes  # noqa: F401
