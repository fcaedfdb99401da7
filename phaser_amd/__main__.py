"""Command line: `python -m phaser_amd run PLAN.yaml` executes every plan in the file on
the MI355X engines (the `phaser run` command, phaser/main.py:12-20); `validate` only
parses and checks the plans; `--seed` makes the whole run reproducible."""
import argparse
import logging
import sys


def main(argv=None):
    ap = argparse.ArgumentParser(prog='python -m phaser_amd')
    sub = ap.add_subparsers(dest='cmd', required=True)
    r = sub.add_parser('run', help='execute a reconstruction plan file')
    r.add_argument('path')
    r.add_argument('--seed', default=None, help='seed for noise, the random object and the groupings')
    v = sub.add_parser('validate', help='parse and check a plan file')
    v.add_argument('path')
    args = ap.parse_args(argv)
    from .plan import plans_from_yaml
    plans = plans_from_yaml(args.path)
    if args.cmd == 'validate':
        for p in plans:
            print(f"{p.name}: {len(p.engines)} engine(s), {sum(e.niter for e in p.engines)} iterations")
        return 0
    logging.basicConfig(level=logging.INFO)
    from .execute import execute_plan
    for p in plans:
        execute_plan(p, seed=args.seed)
    return 0


if __name__ == '__main__':
    sys.exit(main())
