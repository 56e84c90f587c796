"""lible_amd: MI355X-native SHARK electron-repulsion integrals (eri4/eri3/eri2, Schwarz-screened
J/K) behind a C-ABI (include/lbx.h).  `lible_amd.lbx` is the Python binding used by tests and
benchmarks; `lible_amd.basis` builds the basis input."""
