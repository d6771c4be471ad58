"""CPU oracle package (test infrastructure only; see sbf_oracle.h)."""
