"""CPU oracle of the TreePM gravity step — test infrastructure only (see lagpm_oracle.h)."""
