"""MI355X visibility-edge engine: HIP kernels behind a C ABI (libvacd.so),
a ctypes binding (``visacd_amd.native``) and the host driver module
(``visacd_amd._visacd``) behind the drop-in ``visacd`` package."""


def share_hip_runtime():
    """Load torch's HIP runtime before libvacd when torch is installed.

    torch ships its own libamdhip64 (soname libamdhip64.so.7) and links it
    by the file name ``libamdhip64.so``.  If libvacd loads first, the process
    ends up with two HIP/HSA runtimes and the second one sees no GPU.  With
    torch loaded first, libvacd's dependency resolves to torch's copy by
    soname, so both share one runtime, one device list and one stream space.
    """
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
