"""Train the learned far field on the GPU and write the weights blob.

    python -m radiancecascade_amd.train --scenes 10000 --out weights/surrogate.bin

Rows come from the plain cascades on seeded box scenes
(rc_scene_generate(scene_seed0 + s, n_boxes, n_emissive)) at the training
resolution; the network is trained by the library's own forward / backward /
Adam kernels (rc_surrogate_train).  The blob gets a `.json` sidecar with its
SHA-256, the configuration and the per-epoch losses.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--width", type=int, default=512)
    ap.add_argument("--height", type=int, default=512)
    ap.add_argument("--levels", type=int, default=6)
    ap.add_argument("--scene-seed0", type=int, default=10000)
    ap.add_argument("--scenes", type=int, default=10000)
    ap.add_argument("--boxes", type=int, default=32)
    ap.add_argument("--emissive", type=int, default=4)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--batch", type=int, default=16384)
    ap.add_argument("--lr", type=float, default=2e-3)
    ap.add_argument("--lr-decay", type=float, default=0.7)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--out", required=True)
    a = ap.parse_args(argv)

    from . import RadianceCascadeStage
    t0 = time.time()
    with RadianceCascadeStage(64, 64, levels=3) as st:
        print(f"training on {a.scenes} scenes at {a.width}x{a.height} L={a.levels}, {a.epochs} epochs", flush=True)
        rep = st.train_surrogate(width=a.width, height=a.height, levels=a.levels, scene_seed0=a.scene_seed0,
                                 scenes=a.scenes, n_boxes=a.boxes, n_emissive=a.emissive, epochs=a.epochs,
                                 batch=a.batch, lr=a.lr, lr_decay=a.lr_decay, seed=a.seed)
        for e, v in enumerate(rep["epoch_loss"]):
            print(f"epoch {e}: loss {v:.6f}", flush=True)
        meta = {"train": {"width": a.width, "height": a.height, "levels": a.levels, "scene_seed0": a.scene_seed0,
                          "scenes": a.scenes, "n_boxes": a.boxes, "n_emissive": a.emissive, "epochs": a.epochs,
                          "batch": a.batch, "lr": a.lr, "lr_decay": a.lr_decay, "seed": a.seed},
                "report": rep}
        os.makedirs(os.path.dirname(os.path.abspath(a.out)), exist_ok=True)
        sha = st.save_weights(a.out, meta)
    print(json.dumps({"out": a.out, "sha256": sha, "samples": rep["samples"], "steps": rep["steps"],
                      "gen_ms": round(rep["gen_ms"], 1), "train_ms": round(rep["train_ms"], 1),
                      "wall_s": round(time.time() - t0, 1)}), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
