"""pulsesuite_amd: MI355X-native coupled SBE-PSTD time step.

Drop-in modules mirroring PulseSuite's Python surface for this path:
``SBEs`` (quantum-wire SBE solver), ``PSTD3D`` (3-D pseudo-spectral Maxwell
propagator), ``cpml``, ``fftw`` (3-D FFTs), ``typespace``/``typetime``/
``typepulse`` (grid records) and ``rhoPJ`` (wire <-> grid coupling).  All
compute runs in libpss.so (hand-written HIP for gfx950) through a C ABI
(include/pss.h); importing a compute module without the built library raises.
"""

__version__ = "0.1.0"
