"""cylag_amd: MI355X-native particle time step behind CyLag's Python API.

Drop-in names: TriangularMesh, EulerianFieldSet, LagrangianParticleSet,
Parameters, SimuTime, Solver, read_cli and the PCG32 module functions
(cylag/__init__.py, cylag/core/__init__.py, cylag/geom/__init__.py,
cylag/lsm/__init__.py).  The per-particle step runs in libcylag_hip.so.
"""
__version__ = "1.0"

from .constants import *  # noqa: F401,F403
from .cylag_dico import CYLAG_DEFAULT_PARAMETERS  # noqa: F401
from .parameters import Parameters  # noqa: F401
from .simutime import SimuTime  # noqa: F401
from .mesh import TriangularMesh, read_cli  # noqa: F401
from .fields import EulerianFieldSet  # noqa: F401
from .particles import LagrangianParticleSet  # noqa: F401
from .solver import Solver  # noqa: F401
from .batch import BatchSolver  # noqa: F401
from .serve import Request, ServeLoop, latency_stats  # noqa: F401
from .tune import autotune  # noqa: F401
from .sources import (AlwaysScheduler, CountScheduler, IterationDeltaScheduler, IterationScheduler,  # noqa: F401
                      ParticleSource, ParticleSourceFromField, Scheduler, TimeDeltaScheduler, TimeScheduler,
                      circle, init_positions_2d, init_positions_3d, polygon_area, schedules)
from .telemac import SelafinFile  # noqa: F401
from .particles_io import ParticlesIO  # noqa: F401
from .control_shapes import ControlSection  # noqa: F401
from .write_statistics import prepare_secstats_output, write_secstats_buffer  # noqa: F401
from .random_utils import (pcg32_seed, pcg32_advance, random_uniform, random_gaussian,  # noqa: F401
                           random_gaussian_pair, random_gaussian_ziggurat,
                           random_gaussian_ziggurat_pair, generate_random_pair,
                           generate_random_single)
