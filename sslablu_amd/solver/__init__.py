"""Mirror of solver/ (SimonDirckx/SslabLU) on the MI355X."""
