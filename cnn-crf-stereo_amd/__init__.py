"""MI355X-native CNN-CRF stereo hot path: feature CNN, correlation cost volume and
SlackProp CRF inference as hand-written gfx950 HIP kernels behind a C ABI
(include/ccs.h), with the reference's StereoNet / ColorStereoNet / Stereo
operator API in C++ (host/).  This package holds the library sources, the
ctypes binding and the torch front end used by tests and the benchmark."""
import os

PACKAGE_DIR = os.path.dirname(os.path.abspath(__file__))
__version__ = "0.1.0"
