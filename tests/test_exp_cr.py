"""expf_cr, the exp of the build and the reference-form metric
(covertreex_amd/csrc/exp_cr.hpp): binary32 exp correctly rounded, computed
in FP32 arithmetic.  The oracle's correctly rounded mode (EXP_CR, a
double-evaluated expf) is the reference for it.

- CPU: the routine compiled with g++ agrees with the oracle bit for bit on
  every 61st float of [-110, 89] (37M inputs), on the listed hard inputs
  and their neighbours, and on the special values; the oracle's mode is
  itself pinned to the correctly rounded exp by the exhaustive closest case
  (x = -0x1.d2259ap+3, 2.4e-9 ulp from a midpoint) and by exact values.
- GPU: the device routine equals the oracle on every float in [-110, 89]
  (2.24e9 inputs), through cx_expf_cr_range.
"""

import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import oracle.lib as ol

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(ROOT, "covertreex_amd", "csrc")
LO_BITS = int(np.float32(-110.0).view(np.uint32))   # 0xc2dc0000
HI_BITS = int(np.float32(89.0).view(np.uint32))     # 0x42b20000


def _hard_inputs():
    text = open(os.path.join(CSRC, "exp_cr_table.inc")).read()
    body = re.search(r"kExpCrHardX\[\d+\] = \{([^}]*)\}", text).group(1)
    return np.array([int(v.rstrip("u"), 16) for v in re.findall(r"0x[0-9a-f]+u", body)],
                    np.uint32)


def _domain_sample(stride):
    neg = np.arange(0x80000000, LO_BITS + 1, stride, dtype=np.uint64)
    pos = np.arange(0, HI_BITS + 1, stride, dtype=np.uint64)
    return np.concatenate([neg, pos]).astype(np.uint32)


def _oracle(bits):
    return ol.expf_array(np.asarray(bits, np.uint32).view(np.float32), ol.EXP_CR)


@pytest.fixture(scope="module")
def host_exe(tmp_path_factory):
    if shutil.which("g++") is None:
        pytest.skip("needs g++")
    exe = str(tmp_path_factory.mktemp("exp_cr") / "exp_cr_host")
    flags = ["-mfma"] if "fma" in open("/proc/cpuinfo").read() else []
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", *flags, "-I" + CSRC,
                    os.path.join(HERE, "df_host", "exp_cr_host.cpp"), "-o", exe],
                   check=True, capture_output=True)
    return exe


def _host(exe, bits, tmp_path):
    src, dst = tmp_path / "x.bin", tmp_path / "y.bin"
    bits.astype(np.uint32).tofile(src)
    subprocess.run([exe, str(src), str(dst)], check=True)
    return np.fromfile(dst, np.float32)


def test_oracle_mode_is_correctly_rounded(oracle_lib):
    # exact values and the closest-to-midpoint input of the range (its
    # correctly rounded value from a 40-digit evaluation)
    assert ol.expf(0.0, ol.EXP_CR) == 1.0 and ol.expf(-0.0, ol.EXP_CR) == 1.0
    assert ol.expf(-111.0, ol.EXP_CR) == 0.0 and np.isinf(ol.expf(89.5, ol.EXP_CR))
    x = np.float32(float.fromhex("-0x1.d2259ap+3"))
    from decimal import Decimal, getcontext
    getcontext().prec = 40
    exact = Decimal(float(x)).exp()
    got = ol.expf(float(x), ol.EXP_CR)
    lo, hi = np.nextafter(np.float32(got), np.float32(0)), np.nextafter(np.float32(got), np.float32(1))
    assert abs(Decimal(float(got)) - exact) < abs(Decimal(float(lo)) - exact)
    assert abs(Decimal(float(got)) - exact) < abs(Decimal(float(hi)) - exact)
    # smallest subnormal region and the overflow edge
    assert ol.expf(-103.0, ol.EXP_CR) == np.float32(np.exp(np.float64(-103.0)))
    assert ol.expf(88.0, ol.EXP_CR) == np.float32(np.exp(np.float64(88.0)))


def test_host_build_matches_oracle(oracle_lib, host_exe, tmp_path):
    hard = _hard_inputs()
    assert 1 <= hard.size <= 64
    near = np.concatenate([hard, hard - 1, hard + 1])
    edges = np.array([0x00000000, 0x80000000, 0x00000001, 0x80000001, LO_BITS, LO_BITS - 1,
                      HI_BITS, HI_BITS - 1, 0x42b17218, 0x42b17217, 0xc2aeac50, 0xc2cff1b5,
                      0xc2ce8ed0, 0x3f800000, 0xbf800000], np.uint32)
    bits = np.unique(np.concatenate([_domain_sample(61), near, edges]))
    in_range = (bits <= HI_BITS) | ((bits >= 0x80000000) & (bits <= LO_BITS))
    assert in_range.all()
    got = _host(host_exe, bits, tmp_path)
    want = _oracle(bits)
    bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
    assert bad.size == 0, [(hex(int(bits[i])), got[i], want[i]) for i in bad[:5]]
    # outside [-110, 89]: +0 below, +inf above, NaN passes through
    out = np.array([LO_BITS + 1, 0xc3000000, 0xff800000, HI_BITS + 1, 0x7f800000,
                    0x7fc00000, 0xffc00000], np.uint32)
    got = _host(host_exe, out, tmp_path)
    assert got[0] == 0.0 and got[1] == 0.0 and got[2] == 0.0
    assert np.isinf(got[3]) and np.isinf(got[4]) and got[3] > 0
    assert np.isnan(got[5]) and np.isnan(got[6])
    assert np.array_equal(_oracle(out[:5]).view(np.uint32), got[:5].view(np.uint32))


@pytest.mark.gpu
def test_device_matches_oracle_everywhere(hip):
    from covertreex_amd import _ffi

    lib = _ffi.load()
    chunk = 1 << 27
    buf = np.empty(chunk, np.float32)
    ranges = [(0x80000000, LO_BITS), (0, HI_BITS)]
    for lo, hi in ranges:
        for first in range(lo, hi + 1, chunk):
            count = min(chunk, hi + 1 - first)
            print(f"expf_cr range {first:#010x} + {count}", flush=True)
            _ffi.check(lib.cx_expf_cr_range(first, count, _ffi.ptr(buf)))
            got = buf[:count]
            want = ol.expf_range(first, count, ol.EXP_CR)
            bad = np.flatnonzero(got.view(np.uint32) != want.view(np.uint32))
            assert bad.size == 0, [(hex(first + int(i)), got[i], want[i]) for i in bad[:5]]
