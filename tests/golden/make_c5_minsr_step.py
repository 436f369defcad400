"""Write tests/golden/c5_minsr_heavy_tail_step.npz from a dumped C5 MinSR step.

The dump is one step of a C5 MinSR run (kagome Heisenberg 2x4 cells PBC,
N = 24, complex RBM M = 96, 2048 chains x 8 = 16384 samples, MinSR shift
1e-4, lr 0.01, learned proposal with no local moves between learned moves)
taken on the GPU through the decomposed step nqs_sampler_batch ->
nqs_hip_local_energy_batch_ex -> nqs_hip_minsr_solve (parameter space) ->
nqs_ansatz_apply_update: an .npz with `spins` [16384][24] int32, `er` (the
real local energies), `par` (the parameters the step started from) and
`delta` (the device MinSR solution).  Iteration 563 of that run is the step
whose |delta| jumped from about 8 to 540.

    python tests/golden/make_c5_minsr_step.py DUMP.npz

The fixture stores the distinct configurations, the sample -> configuration
index, the device local energy of each distinct configuration, the
parameters and the device delta."""
import os
import sys

import numpy as np

OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "c5_minsr_heavy_tail_step.npz")

z = np.load(sys.argv[1])
spins, er, par, delta = z["spins"], z["er"], z["par"], z["delta"]
packed = np.packbits(spins > 0, axis=1)
key = packed.view(np.dtype((np.void, packed.shape[1]))).ravel()
_, first, inv = np.unique(key, return_index=True, return_inverse=True)
configs = spins[first].astype(np.int8)
e_conf = er[first]
# every sample of a configuration carries the same local energy
assert np.array_equal(e_conf[inv], er)
np.savez_compressed(OUT, configs=configs, index=inv.astype(np.uint16), e_loc=e_conf, params=par,
                    delta_device=delta, shift=np.array([1e-4]))
print(f"wrote {OUT}: {configs.shape[0]} configurations, {inv.size} samples")
