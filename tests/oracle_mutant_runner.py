"""Runs named oracle checks (tests/oracle_checks.py) against one oracle
library in a process of its own: exit 0 when every check passes, 1 when one
fails its assertion, 2 on any other error.  test_oracle_mutants.py runs it
on copies of the oracle built with a planted mistake."""
import os
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
sys.path.insert(0, os.path.dirname(HERE))


def main(argv):
    import exocubed_amd as ex
    from oracle import oracle as orc

    import oracle_checks

    L = orc.load(argv[0])
    try:
        for name in argv[1:]:
            getattr(oracle_checks, "check_" + name)(L, orc, ex)
    except AssertionError as e:
        print(f"detected by {name}: {str(e)[:200]}")
        return 1
    return 0


if __name__ == "__main__":
    try:
        sys.exit(main(sys.argv[1:]))
    except SystemExit:
        raise
    except BaseException as e:  # noqa: BLE001 - report, do not count as detected
        print(f"error: {e!r}")
        sys.exit(2)
