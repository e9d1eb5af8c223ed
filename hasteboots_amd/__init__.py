"""MI355X-native TFHE gate-bootstrapping engine (HIP kernels for gfx950 behind a C ABI).

The Python layer mirrors the reference `tfhe` crate's gate API; see tfhe.py.
"""
from .tfhe import (  # noqa: F401
    GATES,
    BABYBEAR_BINARY_128_BITS_PARAMETERS,
    GOLDILOCKS_BINARY_128_BITS_PARAMETERS,
    ZAMA_BINARY_128_BITS_PARAMETERS,
    SET_BABYBEAR,
    SET_GOLDILOCKS,
    SET_ZAMA,
    BatchGraph,
    Decryptor,
    Encryptor,
    Evaluator,
    KeyGen,
    LWECiphertext,
    Parameters,
    SecretKeyPack,
    TRACE_PARTS,
    device_count,
    last_kernel_ms,
    LATENCY_DEFAULT,
    latency_mode,
    latency_modes,
    mixed_batch,
    restore_latency_modes,
    set_latency_mode,
    set_tuned_config,
    tuned_config,
)
