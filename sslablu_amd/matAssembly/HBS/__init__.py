"""Mirror of matAssembly/HBS/ (SimonDirckx/SslabLU) on the MI355X."""
