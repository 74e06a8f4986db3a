"""Image error metrics of the reference's ImageCompare tool.

Source/Tools/ImageCompare/ImageCompare.cpp:191-211: MSE = mean (a-b)^2 and the
relative MSE = mean (a-b)^2 / (a^2 + 1e-3), both over every stored component (RGBA),
with `a` the first (test) image.
"""
import numpy as np


def mse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(((a - b) ** 2).mean())


def rel_mse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(((a - b) ** 2 / (a ** 2 + 1e-3)).mean())
