"""Train the learned candidate scorer on generated instances and save its weights.

    python -m tsp_gpu_clean_amd.train_scorer --instances 2000 --steps 3000 --out scorer.npz [--eval]

Instances: n uniform in [200, 5000], half uniform and half clustered coordinates.  Labels: edges
of the best tour of a batched memetic GA whose survivors are 2-opt'ed to the local optimum on
their 32-NN candidate lists.  Training: minibatch Adam with the device forward/backward kernels.
``--eval`` then times learned top-8 against kNN-20 (and kNN-8) to a 2 % gap from the
full-neighbourhood 2-opt local optimum at n = 10000, P = 256.
"""

from __future__ import annotations

import argparse
import json
import sys
import time


def main(argv=None) -> int:
  ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
  ap.add_argument("--instances", type=int, default=2000)
  ap.add_argument("--steps", type=int, default=3000)
  ap.add_argument("--batch", type=int, default=65536)
  ap.add_argument("--lr", type=float, default=2e-3)
  ap.add_argument("--seed", type=int, default=0)
  ap.add_argument("--population", type=int, default=8)
  ap.add_argument("--generations", type=int, default=3)
  ap.add_argument("--chunk", type=int, default=64)
  ap.add_argument("--out", default="scorer.npz")
  ap.add_argument("--eval", action="store_true")
  ap.add_argument("--eval-n", type=int, default=10000)
  ap.add_argument("--eval-pop", type=int, default=256)
  args = ap.parse_args(argv)

  from . import scorer as S

  t0 = time.perf_counter()

  def progress(done, secs):
    print(f"labelled {done}/{args.instances} instances in {secs:.1f}s", flush=True)

  sc, info = S.train_on_generated(args.instances, args.steps, args.batch, args.lr, args.seed, args.population,
                                  args.generations, args.chunk, progress)
  print(json.dumps(info), flush=True)
  sc.save(args.out, info)
  print(f"saved {args.out} after {time.perf_counter() - t0:.1f}s", flush=True)
  if args.eval:
    res = S.evaluate_gap(n=args.eval_n, P=args.eval_pop, scorer=sc,
                         progress=lambda name, r: print(name, json.dumps(r), flush=True))
    print(json.dumps(res), flush=True)
  return 0


if __name__ == "__main__":
  sys.exit(main())
