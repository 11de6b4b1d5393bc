"""Persisted shape parameters of the batched search, keyed by the graph's
scale and degree histogram (its size and degree skew).

g5_msbfs_autotune picks the shapes for one graph by timing candidate
searches; the winner is kept in tuned_shapes.json next to this file, so that
later runs on a graph of the same scale and degree distribution start from
it instead of tuning again (and never tune on the roots they time).
"""
import json
import os

PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuned_shapes.json")


def graph_key(graph, scale):
    """'s<scale>-h<histogram>': the scale and the rows per log2 degree class."""
    hist = graph.degree_hist()
    last = max((k for k in range(len(hist)) if hist[k]), default=0)
    return f"s{scale}-h" + ".".join(str(int(x)) for x in hist[:last + 1])


def _read():
    try:
        with open(PATH) as f:
            return json.load(f)
    except (OSError, ValueError):
        return {}


def load(key, keys=()):
    """The persisted record for key ({"params": {...}, ...}) or None; None
    too when its params lack one of keys (a record tuned before a shape
    parameter existed is tuned again)."""
    rec = _read().get(key)
    if rec is None or any(k not in rec.get("params", {}) for k in keys):
        return None
    return rec


def save(key, record):
    data = _read()
    data[key] = record
    tmp = PATH + ".tmp"
    with open(tmp, "w") as f:
        json.dump(data, f, indent=1, sort_keys=True)
    os.replace(tmp, PATH)
