"""Accuracy of the double-float routines behind the exact rescore
(covertreex_amd/csrc/df.hpp), compiled for the CPU with g++ and checked
against long double over random arguments: the relative error bounds the
rescore's per-result error bound is built from (df_exp 2^-43, df_div 2^-44,
df_sqrt 2^-43, df_add 2^-44 of the magnitudes added, df_mul 2^-44).  The two
hardware approximations the device code uses (reciprocal and square root,
about 1 ulp) are modelled one ulp off in a direction chosen per call."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_df_routines_accuracy(tmp_path):
    exe = str(tmp_path / "df_host")
    cmd = ["g++", "-O2", "-std=c++17", "-ffp-contract=off",
           "-I" + os.path.join(HERE, "df_host"),
           "-I" + os.path.join(ROOT, "covertreex_amd", "csrc"),
           os.path.join(HERE, "df_host", "df_host.cpp"), "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True)
    out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout
    err = {k: float(v) for k, v in (line.split() for line in out.strip().splitlines())}
    assert err["exp"] <= 2.0 ** -43
    assert err["exp_abs"] <= 1e-40       # results below 1e-30: the low part is subnormal
    assert err["div"] <= 2.0 ** -44
    assert err["sqrt"] <= 2.0 ** -43
    assert err["add"] <= 2.0 ** -44
    assert err["mul"] <= 2.0 ** -44
