"""CPU oracle (test infrastructure only): see oracle.py and fs_oracle.c."""
