"""CPU oracle for the propagation hot path -- TEST INFRASTRUCTURE ONLY.

A from-scratch NumPy/SciPy restatement of solprop's reference backend (``numpy_ref``)
passes and marches for the routes this build runs: the envelope GNLSE with Kerr and
Raman under the embedded Runge-Kutta march, and the modal field grid with Kerr and
photoionisation under the exponential Adams march. Every function cites the reference
file:line it follows.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import anything from here, and only as the checker. The product
(``solprop_amd``) never imports, links or runs this package: its hot path is the HIP
library, and it raises when that library is missing.

Parity status of each part (what pins it) is recorded in the module docstrings:
``passes`` (docstring known answers, closed forms, the Luna plasma witness),
``controller`` and ``rk`` (the reference's Runge-Kutta decision corpus),
``adams`` (phi-functions against augmented matrix exponentials, the Adams decision
corpus where restated).
"""
