"""MI355X-native backend for CP2K's grid task-list path.

collocate:  rho(r) = sum_tasks  P_ab  (r-A)^a (r-B)^b exp(-za|r-A|^2 - zb|r-B|^2)
integrate:  H_ab   = sum_points V(r) (r-A)^a (r-B)^b exp(...), plus forces/virial

The compute lives in libcp2k_grid_mi355x.so (HIP, gfx950), exporting the C-ABI
of CP2K's src/grid/grid_task_list.h.  This package only builds, loads and
feeds it.
"""
from .tasklist import BasisSpec, TaskListArgs, ncoset  # noqa: F401
