"""Write environments.json: the environment blocks the reference ships (the hill asset's own, in
runtime/assets/hill_lite_bin/asset.json, and the named skies runtime/assets/env/{clearsky30,day}.json), copied as
data: SH coefficients, horizon lobes, ground and glow colours, lights. The display tests build their Env and SkyBg
uniforms from this file and never read the reference.
"""
import json
import os

SRC = "/root/reference/runtime/assets"
OUT = os.path.join(os.path.dirname(os.path.abspath(__file__)), "environments.json")


def main():
    out = {"hill": json.load(open(os.path.join(SRC, "hill_lite_bin", "asset.json")))["environment"]}
    for name in ("clearsky30", "day"):
        e = json.load(open(os.path.join(SRC, "env", name + ".json")))
        out[name] = {k: e[k] for k in ("sky", "lights") if k in e}
    for e in out.values():
        e.pop("note", None)
    with open(OUT, "w") as f:
        json.dump(out, f, indent=1)


if __name__ == "__main__":
    main()
