"""CPU oracle of the implicit PNCG-IPC step — TEST INFRASTRUCTURE ONLY.

A float64 numpy restatement of the reference algorithm (Diff_2D_FEM implicit/solver.py,
implicit/kernels.py, core/math.py, core/sdf.py, implicit/ipc_contact.py, implicit/affine.py,
implicit/actuation.py, implicit/integrator.py, implicit/adjoint.py), the reference results the
GPU path is compared against. Only tests/, __graft_entry__.smoke() and the cpu_baseline leg
of bench.py may import it; the product package never does.

Every per-environment reduction is a deterministic in-order sum (numpy add.at / bincount in
thread-id order), which is what Warp's CPU backend computes; the GPU path may sum in another
order, so parity with it is to a stated tolerance, not bitwise.

Pinning (which check fixes each function, so that a dropped term, a wrong sign or a swapped
index fails one of them) is listed in tests/test_oracle.py's module docstring.
"""
