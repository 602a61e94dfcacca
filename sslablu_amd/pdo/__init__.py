"""Mirror of pdo/ (SimonDirckx/SslabLU) on the MI355X."""
