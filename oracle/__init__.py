"""CPU ORACLE -- test infrastructure only.

Two independent CPU restatements of stonewall3d/lbm3d.py's kernels, used to check the
HIP kernels and nothing else.  Only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg import this package; the product path never does.

  oracle.c_oracle : ctypes binding of lbm_oracle.c (op-for-op FP32 / FP64 restatement in
                    the reference layout, OpenMP over x), built by oracle/Makefile.
  oracle.np_lbm   : a vectorised NumPy restatement written from the equations (FP64),
                    for lattices of a few thousand cells.

Parity pins (what checks each oracle function; see tests/test_oracle_*.py):
  tables, feq           closed-form moments of the equilibrium; OPP/MIRZ from C
  init (noise = 0)      closed form feq(1, (uin, 0, 0)) / w_i
  init noise            PARITY UNPINNED against the reference (Warp RNG is unavailable);
                        pinned GPU-vs-oracle only
  step streaming / BCs  tracer populations on tiny lattices (exact)
  step collision        conservation, rest-state fixed point, second-moment closed form,
                        Cs = 0 -> plain regularized BGK, shear-wave decay rate
  step force            zero at rest; single-link closed form
  step (whole)          C FP64 vs the NumPy restatement; C FP32 vs C FP64
  macro / accum         the NumPy restatement (FP64 and FP32); the noise-free initial
                        state's closed form (rho = 1, u = (uin, 0, 0)); accumulation twice
                        = 2 U exactly; 16-bit macro = macro of the decoded populations
  Q-criterion frames    closed forms (rigid rotation Q = w^2, pure strain pooled to 0,
                        zero at and next to solids) and the NumPy restatement from the
                        symmetric / antisymmetric split of the velocity gradient
  FP16 codec            numpy's float16, exhaustively
  BF16 dithered codec   its properties (BF16 values kept, one of the two neighbours, share
                        rounded up = discarded fraction)
"""
