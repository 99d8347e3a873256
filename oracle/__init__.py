"""CPU oracle for the MI355X statevector engine -- TEST INFRASTRUCTURE ONLY.

This package restates, in plain numpy, the reference's golden model of the QSV2 MUX2 stream
(host/sv/model.py, host/sv/isa.py) plus an independent gate-level complex128 simulator and the
reference's judge metrics (tests/functional_suite.py:90-94, bench/fpga/sv_maxwidth_ref.py:15-25).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline`` leg may import it,
and only as the checker (or, for the CPU baseline, as the timed reference CPU path). The product
path (the package's runtime and the HIP library) never imports or calls anything here.

Modules
  isa    -- QSV2 stream constants (host/sv/isa.py:5-32)
  model  -- decode / tile_index_map / apply_op_tile / run_segment / remap / unpermute / run_program
            in complex64 (reference arithmetic) and complex128 (judge) (host/sv/model.py:7-95)
  dense  -- gate-level complex128 simulator built from Pauli exponentials (no MUX2 machinery)
  judge  -- fidelity, 1-F without cancellation, max|da| (FP64)

How each function is pinned (and what stays unpinned) is listed in each module's docstring.
"""
