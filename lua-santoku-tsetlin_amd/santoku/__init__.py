"""santoku namespace (host mirror of the reference's Lua module tree)."""
