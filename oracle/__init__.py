"""CPU oracle for the STEAM field-generation hot path.

TEST INFRASTRUCTURE ONLY.  The oracle is the checker the HIP path is compared
against.  Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s
``cpu_baseline`` leg may import it; the ``steam_amd`` package never imports,
links or calls anything here, and the product fails loudly (rather than falling
back to this code) when its HIP library is missing.

- ``philox_cms.c``  Philox4x32-10, uniforms, CMS (float32 and float64), in C.
- ``model.py``      NumPy restatement of the rest of the path.
- ``build.py``      gcc recipe for ``liboracle.so``.

Parity status per function is listed in ``model.py``'s header; every function
there has at least one independent pin (golden vector, known answer, closed
form or brute force).
"""
