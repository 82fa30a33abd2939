// Reference-named drop-in entry points (placeholder until the branch path lands).
