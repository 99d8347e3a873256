"""QSV2 stream constants, restated from the reference's format spec (host/sv/isa.py:5-32).

Stream: header [MAGIC, L, n_passes, k, 0, 0, 0, 0]; per pass [m, h0, h1, n_ops, h2, h3, 0, 0];
per op [OP_MUX2, t, cmask, 0 x 5, M0 (8 f32: m00 m01 m10 m11 as re, im), M1 (8 f32)].

Pinned by: the literal zero segment of the reference runtime (host/sv/runtime.py:208), which
tests/test_oracle.py decodes with these constants, and by encode/decode round trips.
"""
MAGIC = 0x51535632          # ASCII "QSV2"
HDR_WORDS = 8
PASS_WORDS = 8
OP_WORDS = 24
OP_MUX2 = 1
MAX_OPS_PER_PASS = 64
MAX_HIGH = 4

# This project's extension for the MI355X retarget: a pass with m > 4 high qubits is followed by
# an 8-word record [OP_HIGHX, h4, h5, h6, h7, h8, h9, h10]; reference streams (m <= 4) never hold it.
OP_HIGHX = 0x48
# Fused form of this project: [OP_DENSE, f, q0..q4, 0] + 2 * 4^f float32 words (dense U, row-major re, im).
OP_DENSE = 2
