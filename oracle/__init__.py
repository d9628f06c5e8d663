"""CPU FP64 oracle for the coupled SBE–PSTD time step.

TEST INFRASTRUCTURE ONLY.  This package is a plain NumPy/SciPy restatement of
PulseSuite's algorithms for the quantum-wire SBE step (``SBEs.SBESolver`` class
path), the 3-D PSTD propagator with CPML and soft source, the wire<->grid
coupling glue and the helpers they call.  Every function cites the reference
file:line it follows (paths relative to the reference's ``src/pulsesuite``).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg
may import it, and only as the checker / the timed CPU baseline.  The product
package (``pulsesuite_amd``) never imports it and has no CPU fallback.

Pinning status per function is listed in each module's header:
  * pinned      – checked against closed forms, invariants, brute-force loop
                  transcriptions or library routines in ``tests/test_oracle_*``;
  * golden      – checked against fixtures written from the reference's own
                  importable field-path modules (``tests/golden``);
  * unpinned    – parity unpinned: no reference fixture or closed form fixes the
                  value; only an independent loop transcription checks it.
"""
