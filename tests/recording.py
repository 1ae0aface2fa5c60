"""Numbers a GPU test measured, written as JSON to $DROPLET_RESULTS_DIR/<name>.json
when that variable is set (the summaries in profiles/ are made from them)."""
import json
import os


def record(name, data):
    d = os.environ.get("DROPLET_RESULTS_DIR")
    if d:
        os.makedirs(d, exist_ok=True)
        with open(os.path.join(d, f"{name}.json"), "w") as fh:
            json.dump(data, fh, indent=1)
