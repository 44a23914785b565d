"""Write tests/golden/ref_kernels.npz: inputs and outputs of the reference's
own main/gpu_kernels.cpp (built for the host by `make -C oracle ref`) on the
seeded case of tests/kernel_cases.make_case(seed=11, small grid).

    make -C oracle ref && python tests/golden/make_ref_vectors.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(HERE))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from kernel_cases import ARGS, host_args, make_case  # noqa: E402
from oracle import oracle  # noqa: E402


def main():
    case = make_case(11, nrows=5, ncols=6, npart=70, hx=0.8, hy=0.55)
    out = {}
    for k, v in case.items():
        out[f"in/{k}"] = np.asarray(v)
    for name in sorted(ARGS):
        args, outs = host_args(name, case)
        oracle.ref_kernel(name)(*args)
        for k, v in outs.items():
            out[f"out/{name}/{k}"] = v
    path = os.path.join(HERE, "ref_kernels.npz")
    np.savez_compressed(path, **out)
    print(path, os.path.getsize(path), "bytes")


if __name__ == "__main__":
    main()
