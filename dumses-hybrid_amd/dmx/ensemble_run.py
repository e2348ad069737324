"""The 48-box float32 ensemble (dmx.problems.ensemble: periodic turbulence
boxes of 32^3-128^3) advanced by one tuned batch, for profiling its
kernels: python -m dmx.ensemble_run [steps]."""
import sys

from dmx import abi
from dmx import problems as P


def main(argv):
    steps = int(argv[1]) if len(argv) > 1 else 5
    boxes = P.ensemble(48, precision=1)
    doms = [abi.Domain(pb.params, pb.uin) for pb in boxes]
    b = abi.Batch(doms)
    ms = b.tune()
    print(f"tuned batch: face/edge {b.shape}, trace {b.trace_shape}, {ms:.3f} ms", flush=True)
    b.advance(2, [0.0] * len(doms))
    b.advance(steps)
    print(f"{steps} steps, {b.last_ms():.3f} ms on the device", flush=True)
    b.close()
    for d in doms:
        d.close()
    return 0


if __name__ == "__main__":
    sys.exit(main(sys.argv))
