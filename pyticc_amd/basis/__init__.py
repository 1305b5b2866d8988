"""Channel bases, quadratures and monomer models of the exact-CC path."""
