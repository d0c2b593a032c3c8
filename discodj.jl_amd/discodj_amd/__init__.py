"""discodj_amd: host side of the MI355X particle-mesh N-body step.

Mirrors the reference's Julia interface for the PM path (PMConfig, PMSolver / pm_acceleration!,
run_nbody, bullfrog_ics, stepper_coefficients) on top of the C ABI of libdiscodj_hip.so (include/discodj_hip.h).
"""
from ._lib import DJError, LIB_PATH, header_exports, lib  # noqa: F401
from .adjoint import density_loss_gradient  # noqa: F401
from . import learned, snapshots  # noqa: F401
from .device import DeviceArray  # noqa: F401
from .cosmology import Cosmology, compute_timetables, linear_power_spectrum, stepper_coefficients  # noqa: F401
from .nbody import add_q, bullfrog_ics, nbody_ics_lpt, run_nbody  # noqa: F401
from .pm import PMConfig, PMSolverHIP, fft3d, ngenic_field, ngenic_field_on_device, pm_acceleration  # noqa: F401
