"""Mirror of multislab/ (SimonDirckx/SslabLU) on the MI355X."""
