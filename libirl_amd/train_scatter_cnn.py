"""Trains the scatter CNN of a workload configuration on the GPU.

    python -m libirl_amd.train_scatter_cnn --config c3 --out libirl_amd/weights/scatter_cnn_c3.bin

Pairs (estimate, attenuation map) -> ESSE terms are generated on the GPU by
the library's own ESSE from seeded ellipsoid phantoms (smoothed and noisy
variants, as OSEM estimates are), and the network is trained with the
library's HIP forward/backward kernels and Adam (irlamd_cnn_train).  The
training statistics go to <out>.json next to the weights.
"""
import argparse
import json
import os
import tempfile
import time

from . import synth


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--config", default="c3")
    ap.add_argument("--out", required=True)
    ap.add_argument("--pairs", type=int, default=3200)
    ap.add_argument("--val", type=int, default=100)
    ap.add_argument("--epochs", type=int, default=36)
    ap.add_argument("--batch", type=int, default=4)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--seed", type=int, default=1)
    a = ap.parse_args(argv)
    import libirl_amd as L
    cfg = synth.CONFIGS[a.config]
    krnl = synth.esse_kernel_file(os.path.join(tempfile.mkdtemp(prefix="irlcnn"), "k.krnl"), cfg)
    t0 = time.time()
    net, st = L.ScatterCnn.train(krnl, cfg["n"], cfg["nz"], cfg["pix"], pairs=a.pairs, val=a.val,
                                 epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed)
    net.save(a.out)
    rec = dict(config=a.config, pairs=a.pairs, val=a.val, epochs=a.epochs, batch=a.batch,
               lr=a.lr, seed=a.seed, wall_seconds=time.time() - t0, **st)
    with open(a.out + ".json", "w") as f:
        json.dump(rec, f, indent=1)
    print(json.dumps(rec))


if __name__ == "__main__":
    main()
