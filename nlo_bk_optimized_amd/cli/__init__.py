"""Command-line front ends (S/cli/)."""
