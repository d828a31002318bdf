"""Small drum ensembles shared by the ensemble tests: uneven drums (a few to a few dozen particles) of
radius 25 mm, each turning at its own speed with its own friction and restitution, three particles of
each drum resting on its mantle so that wall contacts exist from the first step."""
import numpy as np

COUNTS = (40, 8, 25, 3, 15)
RPMS = (60.0, 25.0, 90.0, 10.0, 45.0)
MUS = (0.5, 0.2, 0.8, 0.1, 0.35)
ES = (0.5, 0.9, 0.3, 0.6, 0.75)


def small_ensemble(S, counts=COUNTS, rpms=RPMS, mus=MUS, es=ES, seed=5):
    radius = 0.025
    sc = S.scene_drum_ensemble(list(counts), list(rpms), list(mus), list(es), seed=seed, radius=radius, depth=0.016,
                               n_types=3, n_grid=20, n_nodes=96, wall_spacing=5e-4, semi_axes=(1.5e-3, 2.5e-3),
                               exponents=(2.0, 4.0), gap=0.01)
    P = sc.particles
    drum = sc.meta["drum_of"]
    for k, c in enumerate(sc.meta["centers"]):
        idx = np.flatnonzero(drum == k)[:3]
        for i, x in zip(idx, (-6e-3, 0.0, 6e-3)):
            ext = -sc.shapes[P["shape"][i]]["nodes"][:, 2].min()  # lowest node below the centre
            P["quat"][i] = (1.0, 0.0, 0.0, 0.0)
            P["pos"][i] = c + (x, 0.0, -np.sqrt((radius - ext + 2e-5) ** 2 - x * x))
    return sc
