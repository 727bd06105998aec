"""janc_amd: MI355X-native JANC 2-D reacting-Euler step (WENO5 + global LF +
SSP-RK3 + point-implicit H2-air chemistry) as hand-written gfx950 HIP kernels
behind a C ABI (include/janc_step.h), with a host mirror of the reference's
Simulator interface."""
from .mechanism import Mechanism, DEFAULT_MECH  # noqa: F401


def __getattr__(name):
    if name == "Simulator":
        from .solver import Simulator
        return Simulator
    raise AttributeError(name)
