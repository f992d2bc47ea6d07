"""NumPy model of the E3M12 16-bit population code (storage "e3m12s").

A code is a sign bit and a 15-bit magnitude code r; r decodes to

    |x| = bits_to_float32((r << 11) + (117 << 23)) - 2**-10,

so codes 0..4095 cover [0, 2**-10) in steps of 2**-22 and each further block
of 4096 codes covers [2**(k-10) - 2**-10, 2**(k-9) - 2**-10) in steps of
2**(k-22), k = 1..7.  Encoding rounds |x| + 2**-10 (a float32 sum) to the
nearest code, ties to even, and saturates at r = 0x7fff.  The CPU tests pin
this model by the properties that define the format (monotone decode, block
spacings, round trips of every code, nearest-code rounding); the GPU tests
compare the device conversions with it bit for bit.
"""

import numpy as np

K = np.uint32(117 << 23)          # bits of 2**-10
OFFSET = np.float32(2.0 ** -10)
MAX_CODE = 0x7FFF


def decode(codes):
    """int16 / uint16 codes -> float32 values."""
    h = np.asarray(codes).astype(np.uint16).astype(np.uint32)
    mag = ((h & np.uint32(0x7FFF)) << np.uint32(11)) + K
    m = mag.view(np.float32) - OFFSET
    return (m.view(np.uint32) | ((h & np.uint32(0x8000)) << np.uint32(16))).view(np.float32)


def encode(x):
    """float32 values -> uint16 codes (round to nearest, ties to even; saturating)."""
    x = np.asarray(x, dtype=np.float32)
    t = (np.abs(x) + OFFSET).view(np.uint32).astype(np.uint64)
    r = (t + 0x3FF + ((t >> 11) & 1) - int(K)) >> 11
    r = np.minimum(r, MAX_CODE).astype(np.uint32)
    sign = (x.view(np.uint32) >> np.uint32(16)) & np.uint32(0x8000)
    return (r | sign).astype(np.uint16)


def magnitudes():
    """float64 values of the 32768 magnitude codes, in code order."""
    return decode(np.arange(MAX_CODE + 1, dtype=np.uint16)).astype(np.float64)
