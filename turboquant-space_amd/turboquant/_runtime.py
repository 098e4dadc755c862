"""One HIP runtime per process.

PyTorch-ROCm wheels bundle their own libamdhip64 (soname libamdhip64.so.7),
which torch loads by file name. If libtq_hip.so were loaded first it would
bring in the system runtime under the same soname, torch would then load its
bundled copy as well, and the second runtime in the process sees no device.
Loading torch first (when it is installed) makes libtq_hip.so bind, by soname,
to the runtime torch already holds, so tensors and kernels share one device
context.
"""
try:  # pragma: no cover - depends on the environment
    import torch  # noqa: F401
except ImportError:
    pass
