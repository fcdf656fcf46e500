"""drtopksc_amd -- MI355X-native Dr.Top-k (delegate-centric top-k) for gfx950.

The compute path is the HIP library lib/libdrtopk.so behind the C ABI in
include/drtopk.h; this package is a thin ctypes binding over it, using torch only
for device memory and streams.  There is no CPU fallback: every entry point raises
if the library is missing or no GPU is present.
"""
from .api import (  # noqa: F401
    ALGO_BITONIC, ALGO_BUCKET, ALGO_RADIX, FLAG_BETA, FLAG_FILTER, FLAG_SHUFFLE, FLAG_TIMING, FLAGS_DEFAULT,
    TUNE_DEFAULT, U32, F32, Batch, DrTopKError, Params, Row, Stats, Tuning, lib, lib_path, params, sample_delegates,
    select, tuning_default, tuning_from_table, workspace_size, Server, Query, Result, replay_serial, device_to_host,
    workspace_check, testing_set_segment_cap, file_count, file_read, select_static, STATIC_SCRATCH_BYTES, ALPHA_AUTO, ALPHA_TUNED,
)
