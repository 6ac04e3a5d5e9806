"""Writes naive_cases.npz: input/output vectors of the reference's brute-force
search, searchAlgorithms/radiusNaive.py (torchCompactRadius 0.5.5), for the
cases the closed-form checks do not pin (periodic domains that are not
symmetric about 0, where its periodic `mod` differs from the minimum image;
mixed periodicity; every support mode).

Run in a container that holds the reference source at /root/reference; the
file is loaded on its own (it imports only torch) and run on the CPU.  The
fixture is data: positions, supports, domain, mode and the returned (ii, jj).
"""
import importlib.util
import os

import numpy as np
import torch

SRC = "/root/reference/src/torchCompactRadius/searchAlgorithms/radiusNaive.py"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "naive_cases.npz")


def load():
    spec = importlib.util.spec_from_file_location("radius_naive_ref", SRC)
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod.radiusNaive


def main():
    radiusNaive = load()
    rng = np.random.default_rng(2024)
    cases = []
    domains = [([0.0, 0.0, 0.0], [1.0, 1.0, 1.0]), ([-1.0, -1.0, -1.0], [1.0, 1.0, 1.0]),
               ([-0.3, 0.1, -2.0], [0.9, 1.7, 0.5]), ([2.0, -5.0, 0.25], [3.5, -4.0, 1.0])]
    modes = ["gather", "scatter", "symmetric", "superSymmetric"]
    k = 0
    for dim in (1, 2, 3):
        for di, (mn, mx) in enumerate(domains):
            for mode in modes:
                mn_, mx_ = np.array(mn[:dim], np.float32), np.array(mx[:dim], np.float32)
                ext = mx_ - mn_
                nx = int(rng.integers(40, 160))
                ny = nx if mode == "superSymmetric" else int(rng.integers(40, 160))
                # points inside and somewhat outside the domain
                x = (mn_ + rng.uniform(-0.2, 1.2, (nx, dim)) * ext).astype(np.float32)
                y = (mn_ + rng.uniform(-0.2, 1.2, (ny, dim)) * ext).astype(np.float32)
                hbase = float(ext.min()) * 0.35
                hx = (hbase * rng.uniform(0.5, 1.5, nx)).astype(np.float32)
                hy = (hbase * rng.uniform(0.5, 1.5, ny)).astype(np.float32)
                per = rng.random(dim) < 0.7
                if k % 3 == 0:
                    per[:] = True
                k += 1
                # the reference's caller wraps once (util.py:375-387) before
                # radiusNaive wraps again; record the raw inputs and the caller's view
                xt, yt = torch.from_numpy(x), torch.from_numpy(y)
                mnt, mxt, pt = torch.from_numpy(mn_), torch.from_numpy(mx_), torch.from_numpy(per)
                def once(p):
                    return torch.stack([p[:, i] if not per[i] else
                                        torch.remainder(p[:, i] - mnt[i], mxt[i] - mnt[i]) + mnt[i]
                                        for i in range(dim)], dim=1)
                ii, jj = radiusNaive(once(xt), once(yt), torch.from_numpy(hx), torch.from_numpy(hy), pt, mnt, mxt,
                                     mode)
                cases.append(dict(x=x, y=y, hx=hx, hy=hy, mn=mn_, mx=mx_, per=per, mode=mode,
                                  ii=ii.numpy().astype(np.int64), jj=jj.numpy().astype(np.int64)))
    flat = {}
    for i, c in enumerate(cases):
        for key, v in c.items():
            flat[f"c{i}_{key}"] = np.asarray(v)
    flat["n_cases"] = np.array(len(cases))
    np.savez_compressed(OUT, **flat)
    print(f"wrote {len(cases)} cases, {sum(len(c['ii']) for c in cases)} pairs -> {OUT}")


if __name__ == "__main__":
    main()
