"""GPU: what the trained surrogate gains over the method without it.

Exit spectra come from the batched solver (every point's march the plain solver's, to the
bit), sampled at several lengths a run. Trained on the device on most runs, the network
then predicts the held-out runs' exit spectra; its error is set beside the method without
a network -- each held-out point answered by its nearest training point in the scan's
parameters -- and beside taking the launch spectrum as the answer. The network must beat
both on the same held-out points, and screening with it must put the points it ranks
highest above a random sample in the solver's own objective.
"""

import numpy as np
import pytest

from oracle import fno as oracle
from solprop_amd import surrogate as s

pytestmark = pytest.mark.gpu

# Short capillaries, so that a few hundred runs finish in seconds.
SHORT = {"pressure": (0.2, 3.0), "energy": (20e-6, 200e-6), "fwhm": (5e-15, 30e-15),
         "length": (0.05, 0.3)}


def test_network_beats_the_method_without_it():
    print("surrogate gain on held-out runs", flush=True)
    rng = np.random.default_rng(11)
    runs, k = 384, 4
    pts = s.draw(runs, rng, SHORT)
    lengths = np.sort(rng.uniform(*SHORT["length"], (runs, k)), axis=1)
    solved = s.solve(pts[:, :3], lengths, chunk=runs, max_attempts=6000)
    spectra, params, targets, run_of = s.training_rows(pts[:, :3], lengths, solved, SHORT)
    print("rows", len(spectra), "solver s", solved.seconds, flush=True)
    held = run_of % 6 == 0
    shape = s.Shape(width=32, modes=32, layers=4, hidden=64)
    net = s.Network(shape, max_batch=32)
    o = oracle.Shape(in_ch=shape.in_ch, width=shape.width, modes=shape.modes,
                     layers=shape.layers, hidden=shape.hidden, n=shape.n)
    net.set_parameters(oracle.init(o, seed=2))
    net.load_dataset(spectra[~held], params[~held], targets[~held] - spectra[~held])
    losses, ms = net.train(epochs=60, batch=32, lr=2e-3, seed=1)
    pred = s.predict(net, spectra[held], params[held])
    net.close()
    err_net = s.spectral_error(pred, targets[held])
    nn = s.nearest_neighbour(params[~held], targets[~held], params[held])
    err_nn = s.spectral_error(nn, targets[held])
    err_launch = s.spectral_error(spectra[held], targets[held])
    print("loss", losses[:2], losses[-2:], "train ms", ms, flush=True)
    print("median spectral error (decades): network", np.median(err_net), "nearest neighbour",
          np.median(err_nn), "launch", np.median(err_launch), flush=True)
    print("median density error: network", np.median(s.density_error(pred, targets[held])),
          "nearest neighbour", np.median(s.density_error(nn, targets[held])), flush=True)
    assert np.median(err_net) < 0.8 * np.median(err_nn)
    assert np.median(err_net) < 0.5 * np.median(err_launch)
    # Ranking: the network's highest-ranked held-out rows hold more UV, by the solver.
    objective = s.uv_fraction(pred)
    true = s.uv_fraction(targets[held])
    top = np.argsort(-objective)[: max(1, len(objective) // 10)]
    print("true UV share: network's top tenth", true[top].mean(), "all held-out", true.mean(),
          flush=True)
    assert true[top].mean() > true.mean()
