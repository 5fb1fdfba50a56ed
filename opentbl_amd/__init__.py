"""MI355X-native RK3 substep of the OpenTBL zero-pressure-gradient boundary
layer DNS: compact-FD line solves, z-FFT nonlinear products and the
fractional-step pressure Poisson solve as hand-written HIP kernels behind a
C ABI (libopentbl_amd.so), with the whole box resident in HBM."""
from .config import BoxConfig, CONFIGS, small, ensemble_configs  # noqa: F401
