"""Analysis shared by the Langevin-statistics fixture and its GPU test:
observables of a lipid vesicle and block-averaged means."""
import numpy as np


def area_per_lipid(x):
    """4 pi R^2 / N with R the mean distance of the lipids from their centre."""
    R = np.linalg.norm(x - x.mean(0), axis=1).mean()
    return 4 * np.pi * R * R / len(x)


def block_mean(a, blocks=12):
    """Mean and its standard error from `blocks` consecutive block means
    (blocks longer than the correlation time make them independent)."""
    a = np.asarray(a, np.float64)
    m = len(a) // blocks * blocks
    bm = a[:m].reshape(blocks, -1).mean(1)
    return bm.mean(), bm.std(ddof=1) / np.sqrt(blocks)
