"""Python bindings of libencscore_hip, the MI355X measure path of encode-score (CAMBI banding, window energies,
lasting texture, order-statistic selects), over its C ABI (include/encscore.h).

The HIP library is required: importing this package fails when libencscore_hip.so has not been built (run
`__graft_entry__.build()` or `make -C encode-score_amd`).  There is no CPU fallback.
"""
from .escore import (  # noqa: F401
    ES_BATCHED,
    ES_PLAIN,
    HIST_BINS,
    Cambi,
    DeviceArray,
    EsError,
    EncodeDesc,
    Scorer,
    autotune,
    SynthDesc,
    cambi_params,
    copy_kernel,
    device_count,
    hist_quantiles,
    init,
    lib,
    library_path,
    persist_stats,
    select_rank,
    synth_frame,
    synth_frame_host,
    synchronize,
    window_stats,
    window_values,
    Stream,
    Event,
)
