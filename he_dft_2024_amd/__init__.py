"""MI355X-native BCN-4He-DFT time-step hot path (Orsay-Trento mean field,
13-point stencils, RK-Gill / Hamming predictor-corrector) behind the C-ABI in
include/hedft.h.  The HIP library lives in he_dft_2024_amd/_build/libhedft.so;
the Fortran drop-in module is he_dft_2024_amd/fortran/hedft_gpu.f90."""
from .hedft import (Droplet, HedftError, MODE_STATICS, MODE_TDDFT, PREC_FP64, PREC_MIXED,  # noqa: F401
                    batch_run_pc, default_params, grid_supported, job_bytes, lib, make_job,
                    read_density, read_wavefunction, serve, write_density, write_wavefunction)
