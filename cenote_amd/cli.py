"""``python -m cenote_amd.cli render [scene] --spp --width --height --depth
--out`` — the headless render of cenote-cli (crates/cenote-cli/src/main.rs:
30-71, 103-244): accumulate a scene to ``--spp`` samples and write the
linear averages as one multi-layer EXR (beauty, albedo and normal guides,
first-hit depth; output.rs:84-136).

The scene is a ``.ron`` scene file (cenote_amd/scenefile.py; its Settings
fill any flag left out, main.rs:109-132), a ``.pbrt`` scene imported on the
fly (cenote_amd/pbrt.py; its warnings go to stderr), or one of the seeded
built-ins (``cornell``, ``chart`` — the default, the material chart standing
in for the reference's built-in demo — and ``heavy``, with their
configurations' settings).

``python -m cenote_amd.cli import scene.pbrt --out scene.ron`` converts a
pbrt-v4 scene to a scene file (main.rs:246-275): every import warning is
printed, the change-set must apply, and file references are written
relative to the output's directory.
"""

from __future__ import annotations

import argparse
import sys
import time


def _positive(lo, hi=None):
    def parse(text):
        v = int(text)
        if v < lo or (hi is not None and v > hi):
            raise argparse.ArgumentTypeError(f"must be in {lo}..{hi if hi is not None else ''}")
        return v
    return parse


def build_parser() -> argparse.ArgumentParser:
    ap = argparse.ArgumentParser(prog="cenote", description="Cenote: a GPU path tracer (MI355X)")
    sub = ap.add_subparsers(dest="command", required=True)
    r = sub.add_parser("render", help="render a scene to an EXR (the built-in chart when none is named)")
    r.add_argument("scene", nargs="?", default=None,
                   help="a .ron scene file, or cornell, chart or heavy (default: chart)")
    r.add_argument("--spp", type=_positive(1), help="samples per pixel (default: the scene's)")
    r.add_argument("--width", type=_positive(1), help="output width in pixels (default: the scene's)")
    r.add_argument("--height", type=_positive(1), help="output height in pixels (default: the scene's)")
    r.add_argument("--depth", type=_positive(1, 255), help="maximum path length in bounces (default: the scene's)")
    r.add_argument("--out", default="render.exr", help="output EXR path")
    i = sub.add_parser("import", help="convert a pbrt-v4 scene to a .ron scene file")
    i.add_argument("scene", help="the pbrt-v4 scene file")
    i.add_argument("--out", required=True, help="output .ron path (generated assets are written beside it)")
    return ap


def render(args) -> int:
    import cenote_amd as cn
    from cenote_amd import output, workloads

    from cenote_amd import scenefile

    name = args.scene or "chart"
    if name.endswith(".ron") or name.endswith(".pbrt"):
        try:
            if name.endswith(".pbrt"):
                from cenote_amd import pbrt

                ops, warnings = pbrt.import_scene(name)
                for w in warnings:
                    print(f"warning: {w}", file=sys.stderr)
                desc = scenefile.Description()
                desc.apply(ops)
                data, settings = scenefile.lower(desc)
            else:
                data, settings = scenefile.load_scene(name)
        except (OSError, scenefile.SceneFileError) as e:
            print(f"error: loading scene {name}: {e}", file=sys.stderr)
            return 2
        cfg = {"width": settings["resolution"][0], "height": settings["resolution"][1], "spp": settings["spp"],
               "max_bounces": settings["max_bounces"]}
    elif name in workloads.CONFIGS:
        data, cfg = None, workloads.CONFIGS[name]
    else:
        print(f"error: unknown scene {name!r} (a .ron or .pbrt scene file, or a built-in: "
              f"{', '.join(sorted(workloads.CONFIGS))})", file=sys.stderr)
        return 2
    width = args.width or cfg["width"]
    height = args.height or cfg["height"]
    spp = args.spp or cfg["spp"]
    depth = args.depth or cfg["max_bounces"]
    if not 1 <= depth <= 255:
        print(f"error: max bounces {depth} outside 1..255", file=sys.stderr)
        return 2
    ctx = cn.Context(0)
    scene = cn.Scene(ctx, data if data is not None else workloads.build(name))
    renderer = cn.Renderer(ctx, max_bounces=depth)
    film = cn.Film(ctx, width, height)
    t0 = time.perf_counter()
    renderer.accumulate(scene, film, count=spp)
    beauty, albedo, normal, z = film.averages()
    seconds = time.perf_counter() - t0
    output.write_aov_exr(args.out, beauty, albedo, normal, z)
    print(f"wrote {args.out} ({width}×{height}, {spp} spp; layers: beauty, albedo, normal, Z) "
          f"in {seconds:.2f} s")
    film.close()
    renderer.close()
    scene.close()
    ctx.close()
    return 0


def import_pbrt(args) -> int:
    import os
    from pathlib import Path

    from cenote_amd import pbrt, ron, scenefile

    out = Path(os.path.abspath(args.out))
    try:
        ops, warnings = pbrt.import_scene(args.scene, generated=out.parent)
        for w in warnings:
            print(f"warning: {w}", file=sys.stderr)
        scenefile.Description().apply(ops)  # a dangling reference fails here, not at first render
    except (OSError, scenefile.SceneFileError) as e:
        print(f"error: importing {args.scene}: {e}", file=sys.stderr)
        return 2
    from cenote_amd.ron import Some

    for op in ops:
        for holder, key in scenefile.paths_of(op):
            val = holder[key]
            if isinstance(val, Some):
                holder[key] = Some(os.path.relpath(val.value, out.parent))
            else:
                holder[key] = os.path.relpath(val, out.parent)
    out.write_text(ron.dumps({"version": scenefile.FORMAT_VERSION, "ops": ops}) + "\n")
    print(f"wrote {out} ({len(ops)} ops, {len(warnings)} warnings)")
    return 0


def main(argv=None) -> int:
    args = build_parser().parse_args(argv)
    if args.command == "render":
        return render(args)
    if args.command == "import":
        return import_pbrt(args)
    return 2


if __name__ == "__main__":
    sys.exit(main())
