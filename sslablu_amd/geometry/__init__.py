"""Mirror of geometry/ (SimonDirckx/SslabLU) on the MI355X."""
