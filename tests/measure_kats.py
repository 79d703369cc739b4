"""Known answers of the reference's label and threshold tests
(measure/tests/test_ccomp.py, filters/tests/test_thresholding.py), as data:
(input, keyword arguments, expected labels) triples and the small threshold
image.  Shared by the oracle pins and the GPU tests."""

import numpy as np

BG = 0

X2 = np.array([[0, 0, 3, 2, 1, 9],
               [0, 1, 1, 9, 2, 9],
               [0, 0, 1, 9, 9, 9],
               [3, 1, 1, 5, 3, 0]])
L2 = np.array([[0, 0, 1, 2, 3, 4],
               [0, 5, 5, 4, 2, 4],
               [0, 0, 5, 4, 4, 4],
               [6, 5, 5, 7, 8, 0]])
L2_NOBG = L2 + 1
L2_NOBG[-1, -1] = 10
L2_BG9 = L2_NOBG.copy()
L2_BG9[X2 == 9] = 0
L2_BG9[L2_BG9 > 5] -= 1

X3 = np.zeros((3, 4, 5), int)
X3[0] = [[0, 3, 2, 1, 9], [0, 1, 9, 2, 9], [0, 1, 9, 9, 9], [3, 1, 5, 3, 0]]
X3[1] = [[3, 3, 2, 1, 9], [0, 3, 9, 2, 1], [0, 3, 3, 1, 1], [3, 1, 3, 3, 0]]
X3[2] = [[3, 3, 8, 8, 0], [2, 3, 9, 8, 8], [2, 3, 0, 8, 0], [2, 1, 0, 0, 0]]
L3 = np.zeros((3, 4, 5), int)
L3[0] = [[0, 1, 2, 3, 4], [0, 5, 4, 2, 4], [0, 5, 4, 4, 4], [1, 5, 6, 1, 0]]
L3[1] = [[1, 1, 2, 3, 4], [0, 1, 4, 2, 3], [0, 1, 1, 3, 3], [1, 5, 1, 1, 0]]
L3[2] = [[1, 1, 7, 7, 0], [8, 1, 4, 7, 7], [8, 1, 0, 7, 0], [8, 5, 0, 0, 0]]


def _bg3():
    x = np.zeros((2, 3, 3), int)
    x[0] = [[1, 0, 0], [1, 0, 0], [0, 0, 0]]
    x[1] = [[0, 0, 0], [0, 1, 5], [0, 0, 0]]
    lnb = x.copy()
    lnb[0] = [[1, 2, 2], [1, 2, 2], [2, 2, 2]]
    lnb[1] = [[2, 2, 2], [2, 1, 3], [2, 2, 2]]
    lb = x.copy()
    lb[0] = [[1, BG, BG], [1, BG, BG], [BG, BG, BG]]
    lb[1] = [[BG, BG, BG], [BG, 1, 2], [BG, BG, BG]]
    return x, lb, lnb


def _two3():
    x = np.zeros((2, 3, 3), int)
    x[0] = [[0, 0, 6], [0, 0, 6], [5, 5, 5]]
    x[1] = [[6, 6, 0], [5, 0, 0], [0, 0, 0]]
    lb = x.copy()
    lb[0] = [[BG, BG, 1], [BG, BG, 1], [2, 2, 2]]
    lb[1] = [[1, 1, BG], [2, BG, BG], [BG, BG, BG]]
    return x, lb


def label_cases():
    """[(x, kwargs, expected)] from test_ccomp.py."""
    d = np.array([[0, 0, 1], [0, 1, 0], [1, 0, 0]])
    c48 = np.array([[0, 1], [1, 0]])
    x3d = np.zeros((3, 3, 3), int)
    x3d[0, 2, 2] = x3d[1, 1, 1] = x3d[2, 0, 0] = 1
    x2 = np.zeros((2, 2, 2), int)
    x2[0, 1, 1] = x2[1, 0, 0] = 1
    l4 = x2.copy()
    l4[1, 0, 0] = 2
    bg3, lb3, lnb3 = _bg3()
    two3, ltwo3 = _two3()
    one = np.zeros((3, 3, 3), int)
    one[1, 1, 1] = 1
    cases = [
        (X2, {}, L2), (X2, {"background": 99}, L2_NOBG), (X2, {"background": 9}, L2_BG9),
        (d, {}, d), (c48, {"connectivity": 1}, np.array([[0, 1], [2, 0]])),
        (c48, {"connectivity": 2}, np.array([[0, 1], [1, 0]])),
        (np.array([[1, 0, 0], [1, 1, 5], [0, 0, 0]]), {},
         np.array([[1, 0, 0], [1, 1, 2], [0, 0, 0]])),
        (np.array([[0, 0, 6], [0, 0, 6], [5, 5, 5]]), {"background": 0},
         np.array([[0, 0, 1], [0, 0, 1], [2, 2, 2]])),
        (np.array([[0, 0, 0], [0, 1, 0], [0, 0, 0]]), {"connectivity": 1, "background": 0},
         np.array([[0, 0, 0], [0, 1, 0], [0, 0, 0]])),
        (X3, {}, L3), (x3d, {}, x3d), (x2, {"connectivity": 1}, l4), (x2, {"connectivity": 3}, x2),
        (bg3, {}, lb3), (bg3, {"background": -1}, lnb3), (two3, {"background": 0}, ltwo3),
        (one, {"connectivity": 1, "background": 0}, one),
    ]
    x1 = np.array((0, 1, 2, 2, 1, 1, 0, 0))
    y1 = np.array((0, 1, 2, 2, 3, 3, 0, 0))
    for shp in ((8,), (1, 8), (8, 1), (1, 8, 1), (8, 1, 1), (1, 1, 8)):
        cases.append((x1.reshape(shp), {}, y1.reshape(shp)))
    return cases


NUM_CASES = [(np.array([[1, 0, 6], [0, 0, 6], [5, 5, 5]]), {}, 3),
             (np.array([[1, 0, 6], [0, 0, 6], [5, 5, 5]]), {"background": -1}, 4)]

# test_thresholding.py TestSimpleImage
SIMPLE = np.array([[0, 0, 1, 3, 5], [0, 1, 4, 3, 4], [1, 2, 5, 4, 1], [2, 4, 5, 2, 1],
                   [4, 5, 1, 0, 0]], dtype=int)
