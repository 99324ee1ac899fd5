"""infoli_amd -- MI355X-native InfOli network update.

Python mirror of the C ABI in include/infoli_gpu.h (ctypes over
infoli_amd/lib/libinfoli_gpu.so).  The compute path is the HIP library; this
module only moves arrays across the boundary.  If the library is missing the
import of a compute entry point raises -- there is no CPU fallback.
"""
from .network import (  # noqa: F401
    FP32,
    FP32_KAHAN,
    FP64,
    FP64_DEGREE_DEFAULT,
    EXEC_AUTO,
    EXEC_GRAPH,
    EXEC_PERSISTENT,
    EXEC_STEP,
    GATHER_DIRECT,
    GATHER_LDS,
    GATHER_STAGED,
    GATHER_TILED,
    GATHER_SWEEP,
    GATHER_NETS,
    VARS,
    ERR_ARG,
    ERR_NONFINITE,
    ERR_STATE,
    InfoliError,
    Network,
    Params,
    gen_directed,
    gen_mask,
    gen_normal,
    gen_rand_init,
    gen_symmetric,
    gen_uniform,
    library,
    library_path,
    read_conn_dense,
    read_conn_triplets,
    read_iapp_file,
    read_values,
    tiled_layout_check,
)
from . import workloads  # noqa: F401

__version__ = "0.1.0"
