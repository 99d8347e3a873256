"""Host binary16 conversions used to split dense matrices for the split-FP16 matrix-core path,
checked against numpy's IEEE float16 (round to nearest even) on edge cases and random values."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "u55c-statevector-30q_amd", "csrc")

DRIVER = r"""
#include <cstdio>
#include "svx_half.h"
int main(int argc, char **argv) {
    FILE *fi = std::fopen(argv[1], "rb"), *fo = std::fopen(argv[2], "wb");
    float v;
    while (std::fread(&v, 4, 1, fi) == 1) {
        uint16_t h = svx::f32_to_f16_bits(v);
        float back = svx::f16_bits_to_f32(h);
        std::fwrite(&h, 2, 1, fo);
        std::fwrite(&back, 4, 1, fo);
    }
    std::fclose(fi);
    std::fclose(fo);
    return 0;
}
"""


@pytest.fixture(scope="module")
def conv(tmp_path_factory):
    d = tmp_path_factory.mktemp("half")
    src, exe = d / "drv.cpp", d / "drv"
    src.write_text(DRIVER)
    subprocess.run(["g++", "-O1", "-std=c++17", "-I", HDR, str(src), "-o", str(exe)], check=True)

    def run(vals):
        vals = np.asarray(vals, dtype=np.float32)
        fi, fo = d / "in.bin", d / "out.bin"
        vals.tofile(fi)
        subprocess.run([str(exe), str(fi), str(fo)], check=True)
        raw = np.fromfile(fo, dtype=np.uint8).reshape(-1, 6)
        h = raw[:, :2].copy().view(np.uint16).ravel()
        back = raw[:, 2:].copy().view(np.float32).ravel()
        return h, back
    return run


def test_edge_cases_match_numpy(conv):
    edge = [0.0, -0.0, 1.0, -1.0, 65504.0, 65519.99, 65520.0, 70000.0, -70000.0, np.inf, -np.inf,
            2.0 ** -14, 2.0 ** -24, 2.0 ** -25, 2.0 ** -25 * 1.0001, 3 * 2.0 ** -25, 2.0 ** -26,
            6.1e-5, 5.96e-8, 1e-10, 16384.0, 1.0 + 2.0 ** -11, 1.0 + 3 * 2.0 ** -11, 2047.5, 2048.5]
    h, back = conv(edge)
    ref = np.asarray(edge, dtype=np.float32).astype(np.float16)
    assert np.array_equal(h, ref.view(np.uint16)), [(e, hex(a), hex(b)) for e, a, b in zip(edge, h, ref.view(np.uint16)) if a != b]
    assert np.array_equal(back, ref.astype(np.float32))


def test_random_values_match_numpy(conv):
    rng = np.random.default_rng(0)
    bits = rng.integers(0, 2 ** 32, size=200000, dtype=np.uint64).astype(np.uint32)
    vals = bits.view(np.float32)
    vals = vals[np.isfinite(vals)]
    scaled = (rng.normal(size=100000) * np.exp2(rng.integers(-30, 17, size=100000))).astype(np.float32)
    for v in (vals, scaled):
        h, back = conv(v)
        ref = v.astype(np.float16)
        assert np.array_equal(h, ref.view(np.uint16))
        assert np.array_equal(back, ref.astype(np.float32))


def test_split_of_scaled_unitary_entries(conv):
    """w 2^14 = hi + lo with hi, lo in binary16: |w - (hi + lo) 2^-14| <= 2^-22 |w| + 2^-39."""
    rng = np.random.default_rng(1)
    w = (rng.uniform(-1, 1, 50000) * np.exp2(rng.integers(-20, 1, 50000))).astype(np.float32)
    ws = (w * np.float32(16384.0)).astype(np.float32)
    hbits, hi = conv(ws)
    lbits, lo = conv((ws - hi).astype(np.float32))
    err = np.abs(w.astype(np.float64) - (hi.astype(np.float64) + lo) / 16384.0)
    assert np.all(err <= np.abs(w) * 2.0 ** -22 + 2.0 ** -39)
