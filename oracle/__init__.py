"""CPU oracle (test infrastructure only; never imported by nsrno_amd)."""
