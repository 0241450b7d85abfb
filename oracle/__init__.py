"""CPU oracle of the reference solver (test infrastructure; never imported by fvm_amd)."""
