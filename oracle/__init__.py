"""Oracle: a float64 CPU restatement of AstroPhot's image-model hot path.

This package is TEST INFRASTRUCTURE. It is the checker that the HIP path is
compared against, and the timed CPU baseline in ``bench.py``.  Only
``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s ``cpu_baseline``
leg may import it; the product package ``astrophot_amd`` never does.

Every function cites the reference file:line it restates (paths relative to
the AstroPhot repository root).  Pinning status of each function is listed in
``oracle/hotpath.py``'s module docstring.
"""
