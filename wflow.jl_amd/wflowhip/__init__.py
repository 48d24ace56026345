"""Python host side of the MI355X implementation of the Wflow.jl `sbm` time step.

`Model` mirrors the reference's step interface over the C ABI of libwflow_hip.so
(include/wflow_hip.h): `update_model` is `update_model!` (Wflow/src/sbm_model.jl:60-92),
`run_timestep` is `run_timestep!` (Wflow/src/Wflow.jl:311-322) with the forcing handed in
instead of read from netCDF, `get_value`/`set_value` are the BMI accessors
(Wflow/src/bmi.jl:203-262).

The HIP library is the only compute path: importing works without it, but constructing a
Model raises if the library is missing or no GPU is visible.
"""

from ._abi import Var  # noqa: F401
from .model import Ensemble, Mlp, Model, load_library, library_path, member_catchment  # noqa: F401
