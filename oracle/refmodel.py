"""Test infrastructure: convert a psim JSON model (the schema read by
psim/src/inputManager.cpp:14-115) into the whitespace-separated description
that oracle/ref_driver.cpp reads, so the reference build needs no JSON
library.  Doubles are written with 17 significant digits (exact round trip).
"""
import json
import sys


def _g(v):
    return repr(float(v))


def json_to_text(model: dict) -> str:
    s = model["settings"]
    lines = [
        " ".join([
            "settings",
            str(int(s.get("num_runs", 1))),
            str(int(s["num_measurements"])),
            str(int(s["num_phonons"])),
            _g(s["sim_time"]),
            _g(s["t_eq"]),
            str(int(s["sim_type"])),
            str(int(s.get("step_interval", 0))),
            "1" if s.get("phasor_sim", False) else "0",
        ])
    ]
    mats = model["materials"]
    lines.append(f"materials {len(mats)}")
    for m in mats:
        d, r = m["d_data"], m["r_data"]
        vals = list(d["la_data"]) + list(d["ta_data"]) + [d["max_freq_la"], d["max_freq_ta"]]
        vals += [r["b_l"], r["b_tn"], r["b_tu"], r["b_i"], r["w"]]
        lines.append(m["name"] + " " + " ".join(_g(v) for v in vals))
    sens = model["sensors"]
    lines.append(f"sensors {len(sens)}")
    for x in sens:
        lines.append(f"{int(x['id'])} {x['material']} {_g(x['t_init'])}")
    cells = model["cells"]
    lines.append(f"cells {len(cells)}")
    for c in cells:
        t = c["triangle"]
        pts = [t["p1"]["x"], t["p1"]["y"], t["p2"]["x"], t["p2"]["y"], t["p3"]["x"], t["p3"]["y"]]
        lines.append(" ".join(_g(v) for v in pts) + f" {int(c['sensorID'])} {_g(c['specularity'])}")
    ems = model["emit_surfaces"]
    lines.append(f"emits {len(ems)}")
    for e in ems:
        vals = [e["p1"]["x"], e["p1"]["y"], e["p2"]["x"], e["p2"]["y"], e["temp"], e["duration"], e["start_time"]]
        lines.append(" ".join(_g(v) for v in vals))
    return "\n".join(lines) + "\n"


if __name__ == "__main__":
    with open(sys.argv[1]) as f:
        model = json.load(f)
    with open(sys.argv[2], "w") as f:
        f.write(json_to_text(model))
