import json
from pathlib import Path

GOLDEN = Path(__file__).resolve().parent / "golden"


def load(name: str):
    p = GOLDEN / f"{name}.json"
    if not p.exists():
        return None
    return json.loads(p.read_text())
