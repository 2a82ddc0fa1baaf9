"""fft3dcl_amd: MI355X (gfx950) HIP backend of the FFT3D video denoise path.

Drop-in for the GPU backend of HumbleUser33/FFT3dCL behind its
`FFT3DCL(clip, ...)` call: the C ABI (include/fft3d_hip.h, libfft3d_hip.so),
the VapourSynth API 4 plugin over it (libfft3dcl_vs.so), and a Python host
mirror of the filter for numpy/torch clips (fft3dcl_amd.filter).
"""
from .capi import FFT3DError, Plane, make_params  # noqa: F401
from .filter import FFT3DCL, Instance, GRAY, RGB, YUV  # noqa: F401
