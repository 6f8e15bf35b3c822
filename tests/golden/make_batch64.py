"""Writes batch64_instances.json: the 64-instance table of the batched
wellbore configuration, drawn by intpointsfem_amd.batch.draw_instances(64, seed=0).

    python tests/golden/make_batch64.py
"""
import json
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from intpointsfem_amd.batch import draw_instances  # noqa: E402

if __name__ == "__main__":
    table = draw_instances(64, seed=0)
    with open(os.path.join(HERE, "batch64_instances.json"), "w") as f:
        json.dump({"seed": 0, "generator": "numpy PCG64", "instances": table}, f, indent=1)
    print(f"wrote {len(table)} instances")
