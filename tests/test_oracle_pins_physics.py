"""The physics pins of tests/oracle_checks.py on the built oracle (CPU).
test_oracle_mutants.py shows that each of them fails when a plausible
mistake is planted in the oracle."""
import pytest

import oracle_checks

CHECKS = sorted(n[len("check_"):] for n in dir(oracle_checks) if n.startswith("check_"))


@pytest.mark.parametrize("name", CHECKS)
def test_pin(orc, ex, name):
    getattr(oracle_checks, "check_" + name)(orc.lib(), orc, ex)
