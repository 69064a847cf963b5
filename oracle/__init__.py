"""CPU oracle (test infrastructure): see oracle/sem_oracle.c."""
