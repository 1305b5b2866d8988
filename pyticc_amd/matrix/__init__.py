"""Channel-representation matrices: centrifugal U and the W(R) contraction basis."""
