"""Extract the reference notebooks' recorded outputs into notebook_golden.json.

Reads notebooks/RL_spacecraft.ipynb and notebooks/RL_quadrotor.ipynb of the
reference checkout as JSON text (nothing is executed) and keeps the printed
training losses and the weights after the first gradient step.
Run:  python tests/golden/make_notebook_golden.py /root/reference
"""
import json
import os
import re
import sys


def outputs(nb_path):
    nb = json.load(open(nb_path))
    out = []
    for c in nb["cells"]:
        for o in c.get("outputs", []):
            if "text" in o:
                out.append("".join(o["text"]))
            elif "data" in o and "text/plain" in o["data"]:
                out.append("".join(o["data"]["text/plain"]))
    return "\n".join(out)


def floats_after(text, key):
    m = re.search(re.escape(key) + r"': Array\(\[([^\]]*)\]", text)
    return [float(v) for v in m.group(1).replace("\n", " ").split(",")]


def main(ref):
    sc = outputs(os.path.join(ref, "notebooks", "RL_spacecraft.ipynb"))
    qd = outputs(os.path.join(ref, "notebooks", "RL_quadrotor.ipynb"))
    g = {
        "spacecraft": {
            "losses": [float(v) for v in re.findall(r"reward: (-?[0-9.]+)", sc)],
            "w_x_after_step1": floats_after(sc, "weights_penalization_reference_state_trajectory"),
            "w_u_after_step1": floats_after(sc, "weights_penalization_control_squared"),
        },
        "quadrotor": {
            "losses": [float(v) for v in re.findall(r"reward: (-?[0-9.]+)", qd)],
            "w_x_after_step1": floats_after(qd, "weights_penalization_reference_state_trajectory"),
            "w_u_after_step1": floats_after(qd, "weights_penalization_control_squared"),
            "w_f_after_step1": floats_after(qd, "weights_penalization_final_state"),
        },
    }
    dst = os.path.join(os.path.dirname(os.path.abspath(__file__)), "notebook_golden.json")
    json.dump(g, open(dst, "w"), indent=1)
    print("wrote", dst)


if __name__ == "__main__":
    main(sys.argv[1] if len(sys.argv) > 1 else "/root/reference")
