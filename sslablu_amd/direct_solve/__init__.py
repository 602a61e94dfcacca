"""Mirror of direct_solve/ (SimonDirckx/SslabLU) on the MI355X."""
