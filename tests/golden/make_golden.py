"""Writes tests/golden/golden.json: output hashes of the reference's own CPU
backend, compiled from its sources (oracle/_ref/libisp_ref.so, built by
`make -C oracle ref` where the reference checkout is mounted).

Each case names its input (a synthetic-camera config or a seeded noisy scene,
whose raw bytes are hashed too, pinning the generators) and its settings; the
fixture holds sha256 of the image, histogram and statistics bytes of two
consecutive frames (so the white balance smoothing is covered), and, for tiny
frames, the bytes themselves.

    python tests/golden/make_golden.py
"""
import hashlib
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, os.path.join(ROOT, "holoscan-zozo-modules_amd"))
sys.path.insert(0, os.path.join(ROOT, "tests"))

import golden_cases  # noqa: E402
import oracle_lib  # noqa: E402


def sha(a):
    return hashlib.sha256(a.tobytes()).hexdigest()


def main():
    if not oracle_lib.ref_available():
        sys.exit("oracle/_ref/libisp_ref.so is missing: run `make -C oracle ref` where the reference is mounted")
    out = {"generator": "reference CPU backend (modules/isp/src/cpu_pipeline.cpp) compiled from its sources",
           "cases": {}}
    for name, frame, params, out_format, exposure in golden_cases.cases():
        ref = oracle_lib.RefPipeline(params)
        frames = []
        for _ in range(2):
            r = ref.process(frame, out_format, True, exposure)
            entry = {"image": sha(r.image), "histogram": sha(r.histogram), "stats": sha(r.stats),
                     "stats_values": [float(v) for v in r.stats]}
            if r.image.size <= 512:
                entry["image_bytes"] = r.image.tolist()
            frames.append(entry)
        out["cases"][name] = {"raw": sha(frame.data), "frames": frames}
        print(name, frames[1]["image"][:16])
    with open(os.path.join(HERE, "golden.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)


if __name__ == "__main__":
    main()
