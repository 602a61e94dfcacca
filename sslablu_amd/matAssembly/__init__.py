"""Mirror of matAssembly/ (SimonDirckx/SslabLU) on the MI355X."""
