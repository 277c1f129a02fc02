"""cpmd_amd: MI355X-native plane-wave inner loop for CPMD (vpsi, rhoofr,
orthonormalisation and nonlocal projector contractions) behind a C ABI."""
