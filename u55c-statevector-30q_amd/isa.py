"""QSV2 record-stream constants (format of the reference's host/sv/isa.py:5-32).

Segment: header [MAGIC, L, n_passes, k, 0, 0, 0, 0]
  pass  : [m, h0, h1, n_ops, h2, h3, 0, 0]; tile qubit j < k-m is physical j, k-m+i is h_i
  op    : [OP_MUX2, t, cmask, 0, 0, 0, 0, 0, M0 (8 f32), M1 (8 f32)], t/cmask in tile coordinates;
          a pair (i0, i0 | 2^t) uses M1 iff cmask != 0 and (i0 & cmask) == cmask, else M0.
The same format is accepted by the HIP library (csrc/svx_api.hip); the retargeted MI355X streams
only use wider tiles (k <= 14) and more gathered high qubits (m <= 8).
"""
MAGIC = 0x51535632
HDR_WORDS = 8
PASS_WORDS = 8
OP_WORDS = 24
OP_MUX2 = 1
MAX_OPS_PER_PASS = 64
MAX_HIGH = 4

# MI355X retarget (gfx950: 160 KB LDS per CU).  A c64 tile of 2^13 amplitudes is 64 KiB, so two
# workgroups share a CU; sub-blocks stay >= 2^6 amplitudes (512 B) contiguous with m <= k - 6.
GPU_K = 13
GPU_MIN_SUB = 6
GPU_MAX_HIGH = GPU_K - GPU_MIN_SUB
GPU_MAX_OPS = 64
# Width of the op clusters the MI355X scheduler keeps inside one pass: a 4-qubit cluster is one
# dense unit of the matrix-core engine (16x16 complex, three real 16x16x4 products).
GPU_CLUSTER = 4
