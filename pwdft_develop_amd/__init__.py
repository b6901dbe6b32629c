"""pwdft_develop_amd -- MI355X-native H*psi and eigensolver filter for PWDFT.

The compute path is libpwdft_hip.so (HIP C++ for gfx950) behind the C-ABI in
include/pwdft_hip.h; this package holds its Python test binding (_lib, ops),
the synthetic workload generator and the build recipe.
"""
__all__ = ["workload", "build"]
