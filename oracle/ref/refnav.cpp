// ORACLE -- test infrastructure only.
// Navigation shim over the reference build (filled in with the navigator).
