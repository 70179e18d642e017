"""The library's process rand() stream (csrc/host/prand.hpp) against the C
library's rand(): the reference draws from glibc's generator seeded with 99
(tapsim.cpp:235), so equal seeds must give equal sequences."""
import ctypes

import numpy as np
import pytest

import tapsim_modified_amd as tsg

libc = ctypes.CDLL(None)
libc.rand.restype = ctypes.c_int
libc.srand.argtypes = [ctypes.c_uint]


@pytest.mark.parametrize("seed", [0, 1, 99, 12345, 2**31 - 1, 2**31 + 5, 2**32 - 1])
def test_stream_equals_glibc_rand(seed):
    tsg.srand(seed)
    libc.srand(seed)
    got = tsg.rand(2000)
    want = np.array([libc.rand() for _ in range(2000)], np.int32)
    assert np.array_equal(got, want)


def test_stream_continues_across_calls():
    tsg.srand(7)
    a = np.concatenate([tsg.rand(5), tsg.rand(300)])
    tsg.srand(7)
    assert np.array_equal(a, tsg.rand(305))
