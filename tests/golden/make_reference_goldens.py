"""Writes reference_goldens.json: the reference's own gtest fixture and golden
vectors as data.

Sources (read as text from the reference):
  tests/testcases/function/distributed/FunctionDistributedTest.cpp:8-15
      5x4x3 tensor on a 2x3x1 Cartesian grid, local buffer t[i] = 10*rank + i
  tests/testcases/function/distributed/ttm.cpp:10-17,25-32
      m1 = 0..15 (4x4), m2 = 0..5 (2x3), golden of (t x_1 m1) x_2 m2
  tests/testcases/function/distributed/gram.cpp:12-14
      golden of t_(1) t_(1)^T
The global tensor is assembled from the per-rank rule with the ceil-div block
bounds of src/distribution.cpp:114-141 (mixed-radix rank -> coordinate,
distribution.cpp:79-82).
"""
import json
import os

import numpy as np

TTM1 = [4752.0, 4896.0, 5040.0, 5888.0, 6032.0, 5310.0, 5478.0, 5646.0, 6620.0, 6788.0,
        5868.0, 6060.0, 6252.0, 7352.0, 7544.0, 6426.0, 6642.0, 6858.0, 8084.0, 8300.0,
        6960.0, 7176.0, 7392.0, 8720.0, 8936.0, 7761.0, 8013.0, 8265.0, 9794.0, 10046.0,
        8562.0, 8850.0, 9138.0, 10868.0, 11156.0, 9363.0, 9687.0, 10011.0, 11942.0, 12266.0]
GRAM1 = [1990.0, 2353.0, 4487.0, 7487.0, 2353.0, 2821.0, 5525.0, 9305.0, 4487.0, 5525.0,
         11599.0, 19819.0, 7487.0, 9305.0, 19819.0, 34039.0]


def global_fixture(shape=(5, 4, 3), par=(2, 3, 1)):
    A = np.zeros(shape)
    nranks = int(np.prod(par))
    for rank in range(nranks):
        r, coord = rank, []
        for p in par:
            coord.append(r % p)
            r //= p
        lo = [-(-shape[d] * coord[d] // par[d]) for d in range(3)]
        hi = [-(-shape[d] * (coord[d] + 1) // par[d]) for d in range(3)]
        ls = [hi[d] - lo[d] for d in range(3)]
        blk = (10.0 * rank + np.arange(int(np.prod(ls)))).reshape(ls, order="F")
        A[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = blk
    return A


if __name__ == "__main__":
    A = global_fixture()
    out = {
        "tensor_shape": [5, 4, 3],
        "tensor_colmajor": A.ravel(order="F").tolist(),
        "m1_shape": [4, 4], "m1_colmajor": list(range(16)),
        "m2_shape": [2, 3], "m2_colmajor": list(range(6)),
        "ttm1_shape": [5, 4, 2], "ttm1_colmajor": TTM1,
        "gram1_shape": [4, 4], "gram1_colmajor": GRAM1,
    }
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "reference_goldens.json")
    with open(path, "w") as f:
        json.dump(out, f, indent=1)
    print("wrote", path)
